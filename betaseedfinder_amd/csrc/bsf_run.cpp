// bsf_run.cpp -- run output files of a headless hunt (see bsf_run.hpp) and the hunt itself:
// deterministic world seeds, stage0stream batches on the GPU, the region attempts of the
// records the Java loop would queue, and the run folder the Java search would leave.
#include "bsf_run.hpp"

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>

#include "bsf_host.hpp"

namespace bsf {

namespace {

constexpr const char* kBuildId = "BetaSeedFinder-MI355X-headless";
constexpr const char* kTopHeader =
        "board,rank,seed,blocks,columns,width,depth,footprintArea,fillPercent,avgThickness,minY,maxY,centerX,centerZ,"
        "chunkRadiusUsed,touchesSideBorder\n";
constexpr const char* kCandidateHeader =
        "attempt,seed,stage0,stage0Y88,coarse,p20,upper,high,p19Pass,p19Score,p19Extreme,searchCenterChunkX,"
        "searchCenterChunkZ,matched,status,rawBlocks,verifiedBlocks,rawRadius,verifiedRadius,rawTouchesSide,"
        "verifiedTouchesSide,columns,width,depth,footprintArea,fillPercent,avgThickness,minY,maxY,centerX,centerZ,"
        "minX,maxX,minZ,maxZ,hotScanned,hotRadius,hotBlocks,hotTouchesSide,finalSource\n";

// Shortest round-trip digits of a finite a >= 0: a = 0.digits x 10^e10.
void shortest_digits(double a, std::string& digits, int& e10) {
    char buf[64];
    const auto res = std::to_chars(buf, buf + sizeof(buf), a, std::chars_format::scientific);
    const std::string s(buf, res.ptr);
    const size_t epos = s.find('e');
    digits.clear();
    for (size_t i = 0; i < epos; ++i)
        if (s[i] != '.') digits.push_back(s[i]);
    e10 = std::stoi(s.substr(epos + 1)) + 1;
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
}

const char* tf(bool b) { return b ? "true" : "false"; }

void write_file(const std::string& path, const std::string& text) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw ArgError("cannot write " + path);
    f << text;
    if (!f) throw ArgError("cannot write " + path);
}

bool exists(const std::string& path) {
    struct stat st;
    return ::stat(path.c_str(), &st) == 0;
}

void make_dirs(const std::string& path) {
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part;
        if (!exists(cur) && ::mkdir(cur.c_str(), 0755) != 0 && !exists(cur)) throw ArgError("cannot create " + cur);
        cur += "/";
    }
}

std::string absolute(const std::string& p) {
    if (!p.empty() && p[0] == '/') return p;
    char buf[4096];
    if (!::getcwd(buf, sizeof(buf))) return p;
    std::string cwd(buf);
    std::string rel = p;
    while (rel.rfind("./", 0) == 0) rel = rel.substr(2);
    return rel.empty() || rel == "." ? cwd : cwd + "/" + rel;
}

void copy_file(const std::string& from, const std::string& to) {
    if (!exists(from)) return;
    std::ifstream in(from, std::ios::binary);
    std::ostringstream ss;
    ss << in.rdbuf();
    write_file(to, ss.str());
}

// sanitizeRunLabel (:4300-4307): runs of other characters become '_', at most 40 characters.
std::string sanitize_label(const std::string& label) {
    std::string out;
    bool in_run = false;
    for (char c : label) {
        const bool ok = (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '.' ||
                        c == '_' || c == '-';
        if (ok) {
            out.push_back(c);
            in_run = false;
        } else if (!in_run) {
            out.push_back('_');
            in_run = true;
        }
    }
    if (out.size() > 40) out.resize(40);
    return out.empty() ? "run" : out;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && static_cast<unsigned char>(s[a]) <= ' ') ++a;
    while (b > a && static_cast<unsigned char>(s[b - 1]) <= ' ') --b;
    return s.substr(a, b - a);
}

int64_t now_ms() {
    return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::system_clock::now().time_since_epoch())
            .count();
}

// RUN_TIME_FORMAT "yyyy-MM-dd_HH-mm-ss" of the local time.
std::string stamp_now() {
    const std::time_t t = std::time(nullptr);
    std::tm tm{};
    localtime_r(&t, &tm);
    char buf[32];
    std::strftime(buf, sizeof(buf), "%Y-%m-%d_%H-%M-%S", &tm);
    return buf;
}

// compareLargestBetter / compareFilledFootprintBetter / compareRawFootprintBetter (:5019-5049).
int icmp(int a, int b) { return a < b ? -1 : (a > b ? 1 : 0); }
int dcmp(double a, double b) { return a < b ? -1 : (a > b ? 1 : 0); }  // finite, positive values
int largest_better(const RunRecord& a, const RunRecord& b) {
    if (int c = icmp(a.blocks, b.blocks)) return c;
    if (int c = icmp(a.columns, b.columns)) return c;
    if (int c = icmp(a.area, b.area)) return c;
    return dcmp(a.thick, b.thick);
}
int filled_better(const RunRecord& a, const RunRecord& b) {
    if (int c = icmp(a.columns, b.columns)) return c;
    if (int c = icmp(a.area, b.area)) return c;
    if (int c = icmp(a.blocks, b.blocks)) return c;
    return dcmp(a.fill, b.fill);
}
int raw_better(const RunRecord& a, const RunRecord& b) {
    if (int c = icmp(a.area, b.area)) return c;
    if (int c = icmp(a.columns, b.columns)) return c;
    if (int c = icmp(a.blocks, b.blocks)) return c;
    return dcmp(a.fill, b.fill);
}

// buildBoardText (:5183-5201) with appendResultDetails (:5203-5220).
std::string board_text(const char* title, const char* order, const std::vector<RunRecord>& sorted, int keep) {
    std::string sb;
    sb += std::string(title) + "\n" + order + "\nKeeping top " + std::to_string(keep) + "\n\n";
    for (size_t i = 0; i < sorted.size(); ++i) {
        const RunRecord& r = sorted[i];
        sb += "#" + std::to_string(i + 1) + "\n";
        sb += "seed=" + std::to_string(r.seed) + "\n";
        sb += "blocks=" + std::to_string(r.blocks) + "\n";
        sb += "chunkRadiusUsed=" + std::to_string(r.radius) + "\n";
        sb += std::string("touchesSideBorder=") + tf(r.side) + "\n";
        sb += "columns=" + std::to_string(r.columns) + "\n";
        sb += "filledFootprintColumns=" + std::to_string(r.columns) + "\n";
        sb += "footprint=" + std::to_string(r.width) + "x" + std::to_string(r.depth) + "\n";
        sb += "footprintArea=" + std::to_string(r.area) + "\n";
        sb += "fillPercent=" + java_fixed(r.fill, 2) + "\n";
        sb += "avgThickness=" + java_fixed(r.thick, 2) + "\n";
        sb += "centerX=" + std::to_string(r.center_x) + "\n";
        sb += "centerZ=" + std::to_string(r.center_z) + "\n";
        sb += "boundsX=" + std::to_string(r.min_x) + " to " + std::to_string(r.max_x) + "\n";
        sb += "boundsZ=" + std::to_string(r.min_z) + " to " + std::to_string(r.max_z) + "\n";
        sb += "yRange=" + std::to_string(r.min_y) + " to " + std::to_string(r.max_y) + "\n";
        sb += "preview=" + r.preview + "\n";
        sb += "\n";
    }
    return sb;
}

// appendBoardCsv (:5340-5361).
void board_csv(std::string& sb, const char* board, const std::vector<RunRecord>& sorted) {
    for (size_t i = 0; i < sorted.size(); ++i) {
        const RunRecord& r = sorted[i];
        sb += std::string(board) + "," + std::to_string(i + 1) + "," + std::to_string(r.seed) + "," +
              std::to_string(r.blocks) + "," + std::to_string(r.columns) + "," + std::to_string(r.width) + "," +
              std::to_string(r.depth) + "," + std::to_string(r.area) + "," + java_fixed(r.fill, 4) + "," +
              java_fixed(r.thick, 4) + "," + std::to_string(r.min_y) + "," + std::to_string(r.max_y) + "," +
              std::to_string(r.center_x) + "," + std::to_string(r.center_z) + "," + std::to_string(r.radius) + "," +
              tf(r.side) + "\n";
    }
}

const char* profile_name(int p) {  // getHuntProfileName (:2412-2417)
    return p == BSF_PROFILE_RECORD80 ? "WORLD RECORD 80k+"
           : p == BSF_PROFILE_RECORD60 ? "RECORD 60k+"
           : p == BSF_PROFILE_MEGA     ? "MEGA 30k+"
                                       : "GENERAL";
}

// HuntProfileConfig (gpu_p20_benchmark.cpp:7270-7290) as the Java getters see it
// (getEffectiveP20MinSamples, getEffectiveUpperMinSamples, getEffectiveHighMinSamples,
// getEffectiveP19MinScore, getHunterCoarseThreshold; IslandSearchEngine.java:2425-2487).
struct HuntGates {
    int p20, upper, high;
    double p19;
    int coarse;
};
const HuntGates kHuntGates[4] = {{1, 5, 5, 4.02, 85}, {1, 8, 6, 4.02, 85}, {3, 19, 12, 6.70, 95}, {3, 19, 20, 6.70, 700}};
constexpr int kStage0MinReentry = 2;  // SearchSettings.hunterStage0MinReentrySamples

// getHunterCoarseThreshold (:2437-2442).
int coarse_threshold(const bsf_run_settings& s) {
    if (s.profile == BSF_PROFILE_RECORD60 || s.profile == BSF_PROFILE_RECORD80) return kHuntGates[s.profile].coarse;
    return s.coarse_min > 0 ? s.coarse_min : 180;
}

}  // namespace

std::string java_fixed(double v, int prec) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "Infinity" : "-Infinity";
    const bool neg = std::signbit(v);
    std::string digits;
    int e10 = 0;
    shortest_digits(std::fabs(v), digits, e10);
    const int keep = e10 + prec;  // digits kept, counted from the first significant one
    std::string kept;
    if (keep >= static_cast<int>(digits.size())) {
        kept = digits + std::string(static_cast<size_t>(keep) - digits.size(), '0');
    } else if (keep >= 0) {
        kept = digits.substr(0, static_cast<size_t>(keep));
        if (digits[static_cast<size_t>(keep)] >= '5') {  // half up
            int i = keep - 1;
            while (i >= 0 && kept[static_cast<size_t>(i)] == '9') kept[static_cast<size_t>(i--)] = '0';
            if (i >= 0) {
                ++kept[static_cast<size_t>(i)];
            } else {
                kept.insert(kept.begin(), '1');
                ++e10;
            }
        }
    }
    std::string ip, fp;
    if (kept.empty()) {
        ip = "0";
        fp = std::string(static_cast<size_t>(prec), '0');
    } else if (e10 <= 0) {
        ip = "0";
        fp = std::string(static_cast<size_t>(-e10), '0') + kept;
    } else {
        ip = kept.substr(0, static_cast<size_t>(e10));
        fp = kept.substr(static_cast<size_t>(e10));
    }
    fp.resize(static_cast<size_t>(prec), '0');
    std::string out = neg ? "-" : "";
    out += ip;
    if (prec > 0) out += "." + fp;
    return out;
}

std::string java_double(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "Infinity" : "-Infinity";
    const std::string sign = std::signbit(v) ? "-" : "";
    const double a = std::fabs(v);
    if (a == 0.0) return sign + "0.0";
    std::string digits;
    int e10 = 0;
    shortest_digits(a, digits, e10);
    if (a >= 1e-3 && a < 1e7) {
        std::string ip, fp;
        if (e10 <= 0) {
            ip = "0";
            fp = std::string(static_cast<size_t>(-e10), '0') + digits;
        } else if (e10 >= static_cast<int>(digits.size())) {
            ip = digits + std::string(static_cast<size_t>(e10) - digits.size(), '0');
            fp = "0";
        } else {
            ip = digits.substr(0, static_cast<size_t>(e10));
            fp = digits.substr(static_cast<size_t>(e10));
        }
        return sign + ip + "." + fp;
    }
    const std::string frac = digits.size() > 1 ? digits.substr(1) : "0";
    return sign + digits.substr(0, 1) + "." + frac + "E" + std::to_string(e10 - 1);
}

void JavaHeap::offer(const RunRecord& x) {
    q_.push_back(x);
    size_t k = q_.size() - 1;
    while (k > 0) {  // siftUpUsingComparator
        const size_t parent = (k - 1) >> 1;
        if (cmp_(x, q_[parent]) >= 0) break;
        q_[k] = q_[parent];
        k = parent;
    }
    q_[k] = x;
}

void JavaHeap::poll() {
    const size_t n = q_.size() - 1;
    const RunRecord x = q_[n];
    q_.pop_back();
    if (n == 0) return;
    size_t k = 0;
    const size_t half = n >> 1;
    while (k < half) {  // siftDownUsingComparator
        size_t child = 2 * k + 1;
        size_t right = child + 1;
        if (right < n && cmp_(q_[child], q_[right]) > 0) child = right;
        if (cmp_(x, q_[child]) <= 0) break;
        q_[k] = q_[child];
        k = child;
    }
    q_[k] = x;
}

bool JavaHeap::add(const RunRecord& r, int keep) {
    if (static_cast<int>(q_.size()) < keep) {
        offer(r);
        return true;
    }
    if (cmp_(r, q_[0]) > 0) {
        poll();
        offer(r);
        return true;
    }
    return false;
}

std::vector<RunRecord> JavaHeap::sorted() const {
    std::vector<RunRecord> v = q_;
    std::stable_sort(v.begin(), v.end(), [&](const RunRecord& a, const RunRecord& b) { return cmp_(b, a) < 0; });
    return v;
}

RunWriter::RunWriter(const std::string& root, const std::string& label, const std::string& stamp,
                     const bsf_run_settings& s)
    : root_(root), label_(label), s_(s), top_(largest_better), filled_(filled_better), raw_(raw_better) {
    if (s.top_keep < 1) throw ArgError("topResultsToKeep must be >= 1");
    const std::string runs = root_ + "/runs";
    make_dirs(runs);
    std::string base = stamp;
    const std::string tl = trim(label);
    if (!tl.empty()) base += "_" + sanitize_label(tl);
    std::string cand = runs + "/" + base;
    for (int suffix = 1; exists(cand); ++suffix) {
        char sfx[16];
        std::snprintf(sfx, sizeof(sfx), "_%02d", suffix);
        cand = runs + "/" + base + sfx;
    }
    dir_ = cand;
    make_dirs(dir_);
    write_file(path("top_largest_results.txt"), "Top largest floating islands\nNo results yet.\n");
    write_file(path("top_floating_results.txt"), "Top largest floating islands\nNo results yet.\n");
    write_file(path("top_filled_footprint_results.txt"), "Top filled-footprint floating islands\nNo results yet.\n");
    write_file(path("top_raw_footprint_results.txt"), "Top raw-footprint floating islands\nNo results yet.\n");
    write_file(path("top_island_sideboards.csv"), kTopHeader);
    write_file(root_ + "/latest_run.txt", absolute(dir_) + "\n");
    write_file(path("hunter_candidates.csv"), kCandidateHeader);
}

bool RunWriter::add(const bsf_run_attempt& a, const bsf_region_result& r) {
    ++attempts;
    const bsf_island_hit* raw = r.normal_kept ? &r.normal_hit : nullptr;
    const bsf_island_hit* hot = r.hot_found ? &r.hot_hit : nullptr;
    const bsf_island_hit* ver = r.found ? &r.hit : nullptr;
    const bool final_is_hot = ver && hot && r.hot_won;
    // the resolved match is the verified normal result itself unless the hot scan won
    const bool final_is_border = ver && raw && r.hot_won && !final_is_hot;
    bool unique = false;
    if (ver) {
        char key[256];
        std::snprintf(key, sizeof(key), "%lld:%d:%d:%d:%d:%d:%d:%d:%d", static_cast<long long>(a.seed), ver->min_x,
                      ver->max_x, ver->min_z, ver->max_z, ver->min_y, ver->max_y, ver->blocks, ver->columns);
        unique = seen_.insert(key).second;
        if (unique) {
            ++matches;
            RunRecord rec;
            rec.seed = a.seed;
            rec.blocks = ver->blocks;
            rec.columns = ver->columns;
            rec.center_x = ver->center_x;
            rec.center_z = ver->center_z;
            rec.min_x = ver->min_x;
            rec.max_x = ver->max_x;
            rec.min_z = ver->min_z;
            rec.max_z = ver->max_z;
            rec.min_y = ver->min_y;
            rec.max_y = ver->max_y;
            rec.width = ver->max_x - ver->min_x + 1;
            rec.depth = ver->max_z - ver->min_z + 1;
            rec.area = rec.width * rec.depth;
            rec.fill = 100.0 * rec.columns / static_cast<double>(std::max(1, rec.area));
            rec.thick = rec.blocks / static_cast<double>(std::max(1, rec.columns));
            rec.radius = r.radius_used;
            rec.side = (ver->flags & 1) != 0;
            rec.preview = dir_ + "/floating_seed_" + std::to_string(a.seed) + "_x" + std::to_string(rec.center_x) +
                          "_z" + std::to_string(rec.center_z) + ".png";
            if (top_.add(rec, s_.top_keep)) ++top_updates;
            filled_.add(rec, s_.top_keep);
            raw_.add(rec, s_.top_keep);
            boards_dirty_ = true;
        } else {
            ++duplicates;
        }
    }
    const bool duplicate = ver && !unique;
    const bool matched = ver && !duplicate;
    const char* source = duplicate        ? "duplicate_merged"
                         : final_is_hot    ? "hot_scan"
                         : final_is_border ? "border_verify"
                         : ver             ? "raw"
                                           : "none";
    const bool ver_side = ver && (ver->flags & 1);
    const char* status = duplicate                     ? "duplicate_region_merged"
                         : ver && final_is_hot && ver_side ? "matched_hot_scan_still_touches_side"
                         : ver && final_is_hot             ? "matched_hot_scan"
                         : ver && ver_side                 ? "matched_still_touches_side"
                         : ver                             ? "matched"
                         : raw                             ? "border_rejected_or_invalid"
                         : r.hot_scan_ran                  ? "hot_scan_no_match"
                                                           : "no_match";
    const bsf_island_hit* fin = ver ? ver : raw;
    const int width = fin ? fin->max_x - fin->min_x + 1 : 0;
    const int depth = fin ? fin->max_z - fin->min_z + 1 : 0;
    const int area = width * depth;
    const double fill = fin ? 100.0 * fin->columns / static_cast<double>(std::max(1, area)) : 0.0;
    const double thick = fin ? fin->blocks / static_cast<double>(std::max(1, fin->columns)) : 0.0;
    std::string row;
    row.reserve(320);
    auto num = [&](long long v) {
        row += std::to_string(v);
        row += ',';
    };
    auto str = [&](const std::string& v) {
        row += v;
        row += ',';
    };
    num(a.attempt);
    num(a.seed);
    num(a.high);  // GpuSeedWork.stage0Score = highCount (:1191)
    num(a.high);  // stage0HighScore
    num(a.coarse);
    num(a.p20);
    num(a.upper);
    num(a.high);
    str(tf(a.p19_pass != 0));
    str(java_fixed(a.p19_score, 6));
    str(tf(a.p19_extreme != 0));
    num(a.center_chunk_x);
    num(a.center_chunk_z);
    str(tf(matched));
    str(status);
    num(raw ? raw->blocks : 0);
    num(ver ? ver->blocks : 0);
    num(raw ? r.normal_radius : 0);
    num(ver ? r.radius_used : 0);
    str(tf(raw && (raw->flags & 1)));
    str(tf(ver_side));
    num(fin ? fin->columns : 0);
    num(width);
    num(depth);
    num(area);
    str(java_fixed(fill, 4));
    str(java_fixed(thick, 4));
    num(fin ? fin->min_y : 0);
    num(fin ? fin->max_y : 0);
    num(fin ? fin->center_x : 0);
    num(fin ? fin->center_z : 0);
    num(fin ? fin->min_x : 0);
    num(fin ? fin->max_x : 0);
    num(fin ? fin->min_z : 0);
    num(fin ? fin->max_z : 0);
    str(tf(r.hot_scan_ran != 0));
    num(hot ? r.hot_radius : 0);
    num(hot ? hot->blocks : 0);
    str(tf(hot && (hot->flags & 1)));
    row += source;
    row += '\n';
    csv_ += row;
    if (csv_.size() >= (size_t(1) << 20)) flush_rows();
    return unique;
}

void RunWriter::flush_rows() {
    if (csv_.empty()) return;
    std::ofstream f(path("hunter_candidates.csv"), std::ios::binary | std::ios::app);
    f << csv_;
    if (!f) throw ArgError("cannot write " + path("hunter_candidates.csv"));
    csv_.clear();
}

void RunWriter::write_boards() {
    if (!boards_dirty_) return;
    const int keep = s_.top_keep;
    const std::vector<RunRecord> top = top_.sorted(), filled = filled_.sorted(), raw = raw_.sorted();
    const std::string t = board_text("Top largest floating islands",
                                     "Sorted by blocks, then occupied columns, then footprint area, then thickness",
                                     top, keep);
    write_file(path("top_largest_results.txt"), t);
    write_file(path("top_floating_results.txt"), t);
    write_file(path("top_filled_footprint_results.txt"),
               board_text("Top filled-footprint floating islands",
                          "Sorted by occupied X/Z columns. This is raw footprint area with fill ratio accounted for.",
                          filled, keep));
    write_file(path("top_raw_footprint_results.txt"),
               board_text("Top raw-footprint floating islands",
                          "Sorted by bounding-box footprint area. Fill percentage is shown so sparse boxes are obvious.",
                          raw, keep));
    std::string csv = kTopHeader;
    board_csv(csv, "largest", top);
    board_csv(csv, "filled_footprint", filled);
    board_csv(csv, "raw_footprint", raw);
    write_file(path("top_island_sideboards.csv"), csv);
}

void RunWriter::close(const std::string& status, int64_t checked, int64_t start_ms, int64_t end_ms) {
    flush_rows();
    write_boards();
    write_file(path("performance_profile.txt"), "Performance profiler disabled.\n");
    write_manifest(status, checked, start_ms, end_ms);
    // publishLatestRunCopies (:4483-4510)
    copy_file(path("hunter_candidates.csv"), root_ + "/hunter_candidates.csv");
    copy_file(path("top_largest_results.txt"), root_ + "/top_largest_results.txt");
    copy_file(path("top_largest_results.txt"), root_ + "/top_floating_results.txt");
    copy_file(path("top_filled_footprint_results.txt"), root_ + "/top_filled_footprint_results.txt");
    copy_file(path("top_raw_footprint_results.txt"), root_ + "/top_raw_footprint_results.txt");
    copy_file(path("top_island_sideboards.csv"), root_ + "/top_island_sideboards.csv");
    copy_file(path("manifest.txt"), root_ + "/latest_manifest.txt");
    copy_file(path("performance_profile.txt"), root_ + "/latest_performance_profile.txt");
}

void RunWriter::fail(int64_t checked, int64_t start_ms, int64_t end_ms) {
    flush_rows();
    write_boards();
    write_manifest("ERROR", checked, start_ms, end_ms);
}

void RunWriter::write_manifest(const std::string& status, int64_t checked, int64_t start_ms, int64_t end_ms) {
    const int prof = s_.profile;
    const HuntGates& g = kHuntGates[prof];
    const bsf_island_params& t = s_.thresholds;
    std::string m = "BetaSeedFinder run manifest\n";
    auto kv = [&](const char* k, const std::string& v) { m += std::string(k) + "=" + v + "\n"; };
    kv("build", kBuildId);
    kv("status", status);
    kv("runFolder", absolute(dir_));
    kv("runLabel", label_);
    kv("startEpochMs", std::to_string(start_ms));
    kv("endEpochMs", std::to_string(end_ms));
    if (end_ms > start_ms) kv("elapsedSeconds", java_fixed((end_ms - start_ms) / 1000.0, 3));
    kv("checked", std::to_string(checked));
    kv("matches", std::to_string(matches));
    kv("topUpdates", std::to_string(top_updates));
    kv("checkedUnit", "radius7Regions");
    kv("regionBudget", std::to_string(s_.regions));
    kv("gpuCoverageCentersPerWorld", "8");
    kv("gpuCoverageCenterSpacingChunks", "15");
    kv("gpuProtocol", "ST0R3708");
    kv("gpuTerrainPerlinCacheStates", "66");
    kv("gpuClimatePerlinCacheStates", "10");
    kv("gpuCachedCoarseInitialization", "true");
    kv("gpuCompactLowerThreads", "32");
    kv("gpuCompactUpperThreads", "192");
    kv("gpuSharedUpperYAxis", "true");
    kv("gpuDirectWriteNoise23", "true");
    m += "\n";
    kv("chunkRadius", std::to_string(s_.chunk_radius));
    kv("seedsToCheck", std::to_string(s_.regions));
    kv("gpuStage0Enabled", "true");
    kv("gpuStage0BatchSize", "32768");
    kv("minBlocks", std::to_string(t.min_blocks));
    kv("minColumns", std::to_string(t.min_columns));
    kv("minYForMatch", std::to_string(t.min_y_for_match));
    kv("minWidth", std::to_string(t.min_width));
    kv("minDepth", std::to_string(t.min_depth));
    kv("minAvgThickness", java_double(t.min_avg_thickness));
    kv("topResultsToKeep", std::to_string(s_.top_keep));
    kv("savePreviews", "false");
    kv("featureLoggingEnabled", "false");
    kv("nativeResponseMode", s_.research ? "research" : "lean");
    kv("nativeResponseBytesPerRegion", s_.research ? "34" : "8");  // the Java value for lean
    m += "\n";
    kv("hunterMode", "true");
    kv("huntProfile", std::to_string(prof));
    kv("huntProfileName", profile_name(prof));
    kv("megaMode", tf(prof == BSF_PROFILE_MEGA));
    kv("recordHuntMode", tf(prof == BSF_PROFILE_RECORD60));
    kv("extremeRecordHuntMode", tf(prof == BSF_PROFILE_RECORD80));
    kv("effectiveP20MinPositiveColumns", std::to_string(g.p20));
    kv("effectiveP19MinScore", java_double(g.p19));
    kv("filterResearchEnabled", tf(s_.research != 0));
    kv("megaUpperMinPositiveColumns", std::to_string(g.upper));
    kv("megaHighMinReentryColumns", std::to_string(g.high));
    kv("hunterCoarseMinCells", std::to_string(coarse_threshold(s_)));
    kv("stage0MinReentry", std::to_string(kStage0MinReentry));
    kv("borderVerifyRadii", "10,13,16,20,24");
    kv("borderVerifyMaxRadius", "24");
    kv("dangerousRejectExactScanMinCoarse", std::to_string(coarse_threshold(s_)));
    kv("hotSeedScanEnabled", "true");
    kv("hotSeedCoarseThreshold", "120");
    kv("hotSeedScanRadius", "16");
    m += "\n";
    kv("performanceProfilerEnabled", "false");
    kv("deterministicSeedMode", "true");
    kv("deterministicSeedSequenceSeed", std::to_string(static_cast<int64_t>(s_.sequence_seed)));
    write_file(path("manifest.txt"), m);
}

namespace {

// Entry points without the library lock (the hunt calls the locked ones itself).
template <class F>
int unlocked_call(F&& f) {
    try {
        f();
        return BSF_OK;
    } catch (const ArgError& e) {
        set_error(e.what());
        return BSF_ERR_ARG;
    } catch (const CapacityError& e) {
        set_error(e.what());
        return BSF_ERR_CAPACITY;
    } catch (const HipError& e) {
        set_error(e.what());
        return BSF_ERR_HIP;
    } catch (const std::exception& e) {
        set_error(e.what());
        return BSF_ERR_STATE;
    }
}

void checked_call(int rc) {
    if (rc == BSF_OK) return;
    const std::string msg = bsf_last_error();
    if (rc == BSF_ERR_ARG) throw ArgError(msg);
    if (rc == BSF_ERR_CAPACITY) throw CapacityError(msg);
    if (rc == BSF_ERR_HIP) throw HipError(msg);
    throw std::runtime_error(msg);
}

void validate_settings(const bsf_run_settings* s) {
    if (!s) throw ArgError("null run settings");
    if (s->profile < 0 || s->profile > 3) throw ArgError("hunt profile must be general, mega, record60, or record80");
    if (s->top_keep < 1) throw ArgError("topResultsToKeep must be >= 1");
    if (s->chunk_radius < 0 || s->chunk_radius > 24) throw ArgError("chunk radius must be 0..24");
    if (s->regions < 0) throw ArgError("negative region budget");
}

}  // namespace
}  // namespace bsf

using namespace bsf;

struct bsf_run {
    RunWriter w;
    int64_t start_ms;
};

extern "C" {

int bsf_run_open(const char* output_root, const char* run_label, const char* stamp, int64_t start_ms,
                 const bsf_run_settings* s, bsf_run** out) {
    return unlocked_call([&] {
        if (!output_root || !out) throw ArgError("null run argument");
        validate_settings(s);
        std::unique_ptr<bsf_run> r(new bsf_run{
            RunWriter(output_root, run_label ? run_label : "", stamp ? stamp : stamp_now(), *s),
            start_ms > 0 ? start_ms : now_ms()});
        r->w.write_manifest("RUNNING", 0, r->start_ms, 0);  // at search start (:761)
        *out = r.release();
    });
}

int bsf_run_add(bsf_run* r, const bsf_run_attempt* a, const bsf_region_result* res, int32_t* unique_match) {
    return unlocked_call([&] {
        if (!r || !a || !res) throw ArgError("null run argument");
        const bool u = r->w.add(*a, *res);
        if (unique_match) *unique_match = u ? 1 : 0;
    });
}

int bsf_run_dir(const bsf_run* r, char* buf, int64_t cap) {
    return unlocked_call([&] {
        if (!r || !buf || cap < 1) throw ArgError("null run argument");
        const std::string& d = r->w.dir();
        if (static_cast<int64_t>(d.size()) >= cap) throw CapacityError("run folder path longer than the buffer");
        std::memcpy(buf, d.c_str(), d.size() + 1);
    });
}

int bsf_run_close(bsf_run* r, const char* status, int64_t checked, int64_t end_ms) {
    return unlocked_call([&] {
        if (!r) throw ArgError("null run argument");
        std::unique_ptr<bsf_run> owned(r);
        const std::string st = status ? status : "FINISHED";
        const int64_t end = end_ms > 0 ? end_ms : now_ms();
        if (st == "ERROR")
            owned->w.fail(checked, owned->start_ms, end);
        else
            owned->w.close(st, checked, owned->start_ms, end);
    });
}

int bsf_hunt(const char* output_root, const char* run_label, const char* stamp, const bsf_run_settings* s,
             bsf_hunt_summary* out, char* run_dir, int64_t run_dir_cap) {
    return unlocked_call([&] {
        if (!output_root) throw ArgError("null run argument");
        validate_settings(s);
        const int64_t t0 = now_ms();
        const auto c0 = std::chrono::steady_clock::now();
        RunWriter w(output_root, run_label ? run_label : "", stamp ? stamp : stamp_now(), *s);
        w.write_manifest("RUNNING", 0, t0, 0);  // at search start (:761)
        const HuntGates& g = kHuntGates[s->profile];
        const int coarse_min = coarse_threshold(*s);
        const int centers = 8, batch = 32768;
        const size_t rb = s->research ? 34 : 16;
        std::vector<int64_t> seeds(batch);
        std::vector<uint8_t> recs(static_cast<size_t>(batch) * centers * rb);
        int64_t issued_regions = 0, issued_worlds = 0, survivors = 0;
        static const int kCx[8] = {0, 15, 0, 15, -15, -15, 0, 15};
        static const int kCz[8] = {0, 0, 15, 15, 0, 15, -15, -15};
        // Survivors wait for a region call until kFlushItems are pending or kFlushBatches
        // batches have passed (a region call over a few windows costs about what one over
        // hundreds costs); they reach the writer in attempt order either way.
        constexpr size_t kFlushItems = 1024;
        constexpr int kFlushBatches = 8;
        std::vector<bsf_region_item> items;
        std::vector<bsf_run_attempt> atts;
        int64_t written_regions = 0, pending_regions = 0;
        int pending_batches = 0;
        auto drain = [&] {
            std::vector<bsf_region_result> res(items.size());
            if (!items.empty())
                checked_call(bsf_region_attempts(items.data(), static_cast<int64_t>(items.size()), &s->thresholds,
                                                 s->chunk_radius, res.data()));
            for (size_t k = 0; k < items.size(); ++k) w.add(atts[k], res[k]);
            survivors += static_cast<int64_t>(items.size());
            items.clear();
            atts.clear();
            written_regions += pending_regions;
            pending_regions = 0;
            pending_batches = 0;
        };
        // a failed call leaves the files of the regions before it and an ERROR manifest (:925-934)
        try {
            while (issued_regions < s->regions) {
                const int64_t regions = std::min<int64_t>(int64_t(batch) * centers, s->regions - issued_regions);
                const int count = static_cast<int>((regions + centers - 1) / centers);
                // world attempts are 1-based (runGpuStage0Pipeline, :1321-1325)
                checked_call(bsf_deterministic_seeds(s->sequence_seed, issued_worlds + 1, count, seeds.data()));
                bsf_stage0_counters ctr{};
                checked_call(bsf_stage0_records(seeds.data(), count, s->profile, s->research, centers, recs.data(), &ctr));
                for (int64_t off = 0; off < regions; ++off) {
                    const uint8_t* r = recs.data() + static_cast<size_t>(off) * rb;
                    bsf_run_attempt a{};
                    a.attempt = issued_regions + off + 1;
                    a.seed = seeds[static_cast<size_t>(off / centers)];
                    a.center_chunk_x = kCx[off % centers];
                    a.center_chunk_z = kCz[off % centers];
                    a.p20 = r[0];
                    a.upper = r[1];
                    a.high = r[2];
                    a.p19_pass = r[3] != 0;
                    a.coarse = r[4] | (r[5] << 8);
                    a.p19_extreme = r[6] != 0;
                    const bool mega = r[7] != 0;
                    uint64_t sb = 0;
                    for (int k = 0; k < 8; ++k) sb |= static_cast<uint64_t>(r[8 + k]) << (8 * k);
                    std::memcpy(&a.p19_score, &sb, 8);
                    // rejectStage 0 (:1366-1376)
                    const bool pass = a.p20 >= g.p20 && a.upper >= g.upper && a.high >= kStage0MinReentry &&
                                      a.high >= g.high && a.p19_pass && !mega && a.coarse >= coarse_min;
                    if (!pass) continue;
                    atts.push_back(a);
                    items.push_back(bsf_region_item{a.seed, a.center_chunk_x, a.center_chunk_z, a.coarse, 0});
                }
                issued_worlds += count;
                issued_regions += regions;
                pending_regions += regions;
                if (items.size() >= kFlushItems || ++pending_batches >= kFlushBatches) drain();
            }
            drain();
        } catch (...) {
            w.fail(written_regions, t0, now_ms());
            throw;
        }
        w.close("FINISHED", issued_regions, t0, now_ms());
        if (out) {
            out->worlds = issued_worlds;
            out->regions = issued_regions;
            out->survivors = survivors;
            out->matches = w.matches;
            out->duplicates = w.duplicates;
            out->top_updates = w.top_updates;
            out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
        }
        if (run_dir && run_dir_cap > 0) {
            if (static_cast<int64_t>(w.dir().size()) >= run_dir_cap)
                throw CapacityError("run folder path longer than the buffer");
            std::memcpy(run_dir, w.dir().c_str(), w.dir().size() + 1);
        }
    });
}

}  // extern "C"
