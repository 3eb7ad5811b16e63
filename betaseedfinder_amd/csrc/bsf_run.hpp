// bsf_run.hpp -- the run output files of a headless hunt, in the formats the Java search writes
// them (IslandSearchEngine.java): the run folder and its initial files (initializeRunOutput,
// :4230-4299), one hunter_candidates.csv row per region attempt (appendCandidateCsv,
// :4826-4982), the duplicate-match filter (registerUniqueMatch, :2062-2070), the three
// leaderboards and their text and CSV renderings (:4985-5361), the manifest (writeManifest,
// :4309-4449) and the copies of the latest run in the output root (publishLatestRunCopies,
// :4483-4510).  Host code only.
#pragma once

#include <cstdint>
#include <functional>
#include <set>
#include <string>
#include <vector>

#include "../../include/bsf.h"

namespace bsf {

// java.util.Formatter "%.<prec>f" with Locale.US: the shortest round-trip decimal digits of v
// (what FloatingDecimal hands the formatter), rounded half up at prec decimals; "NaN",
// "Infinity", and a '-' for every negative value and -0.0 (Double.compare(v, 0.0) < 0).
std::string java_fixed(double v, int prec);
// Double.toString: shortest round-trip digits, plain notation for 1e-3 <= |v| < 1e7 (at least
// one fraction digit), computerized scientific notation ("1.0E7") otherwise.
std::string java_double(double v);

// ResultRecord (:6592-6650) of one match.
struct RunRecord {
    int64_t seed = 0;
    int blocks = 0, columns = 0, center_x = 0, center_z = 0;
    int min_x = 0, max_x = 0, min_z = 0, max_z = 0, min_y = 0, max_y = 0;
    int width = 0, depth = 0, area = 0;
    double fill = 0.0, thick = 0.0;
    int radius = 0;
    bool side = false;
    std::string preview;
};

// java.util.PriorityQueue over a comparator "better(a, b) > 0": a binary min-heap of the worse
// records (siftUp on offer, last-to-root siftDown on poll), so ties keep Java's order.
class JavaHeap {
   public:
    using Cmp = std::function<int(const RunRecord&, const RunRecord&)>;
    explicit JavaHeap(Cmp cmp) : cmp_(std::move(cmp)) {}
    // addToBoard (:5060-5075): true when the record entered the board.
    bool add(const RunRecord& r, int keep);
    // The board best first: a stable sort of the heap array (List.sort is stable).
    std::vector<RunRecord> sorted() const;
    size_t size() const { return q_.size(); }

   private:
    void offer(const RunRecord& r);
    void poll();
    Cmp cmp_;
    std::vector<RunRecord> q_;
};

class RunWriter {
   public:
    // Creates <root>/runs/<stamp>[_<label>] (with _01, _02 ... when taken) and its initial files.
    RunWriter(const std::string& root, const std::string& label, const std::string& stamp, const bsf_run_settings& s);
    // One region attempt, in attempt order: the candidate row, and for a match the duplicate
    // filter and the leaderboards.  Returns whether it was a new (unique) match.
    bool add(const bsf_run_attempt& a, const bsf_region_result& r);
    // Final leaderboard files, the manifest, the performance file and the latest-run copies.
    void close(const std::string& status, int64_t checked, int64_t start_ms, int64_t end_ms);
    // A failed run: the candidate rows and leaderboards so far and an ERROR manifest, no
    // latest-run copies (search(), :925-934).
    void fail(int64_t checked, int64_t start_ms, int64_t end_ms);
    // manifest.txt (writeManifest, :4309-4449); status RUNNING at the start (end_ms 0).
    void write_manifest(const std::string& status, int64_t checked, int64_t start_ms, int64_t end_ms);
    const std::string& dir() const { return dir_; }

    int64_t matches = 0, duplicates = 0, top_updates = 0, attempts = 0;

   private:
    std::string path(const char* name) const { return dir_ + "/" + name; }
    void write_boards();
    void flush_rows();  // appends the buffered candidate rows to hunter_candidates.csv
    std::string root_, label_, dir_;
    bsf_run_settings s_;
    std::set<std::string> seen_;
    JavaHeap top_, filled_, raw_;
    std::string csv_;  // candidate rows not yet appended to hunter_candidates.csv
    bool boards_dirty_ = false;
};

}  // namespace bsf
