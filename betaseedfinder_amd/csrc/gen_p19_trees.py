#!/usr/bin/env python3
"""Regenerate p19_trees.inc, the P19 "monster gate" tree ensemble as a flat node table.

The gate is an empirically trained model: 16 regression trees (depth <= 5) over
Stage0 topology features, score = (tree0 + ... + tree15) / 16
(native/src/p19_monster_gate.hpp:48-67, src/P19MonsterGate.java:18-49).  Its
split thresholds and leaf values are data that must stay bit-identical, so this
script reads them out of the reference header once and writes them as a table;
the evaluator that walks the table is written independently in this project.

Usage (needs the reference checkout; the output is committed):
    python3 gen_p19_trees.py /root/reference/native/src/p19_monster_gate.hpp > p19_trees.inc
"""
import re
import sys

# Feature expressions as they appear in the reference, mapped to this project's feature ids.
FEATURES = [
    "((double) p.stage0FullY88 / (scout + EPSILON))",
    "p.stage0FullY88",
    "p.stage0FullY96",
    "(p.stage0Y88Width + p.stage0Y88Depth + p.stage0Y96Width + p.stage0Y96Depth)",
    "(p.stage0Y96Width * p.stage0Y96Depth)",
    "scout",
    "(p.stage0Y88Width * p.stage0Y88Depth)",
    "((double) p.stage0Y88LargestCluster / (p.stage0FullY88 + EPSILON))",
    "p.stage0Y96Width",
    "(p.stage0FullY96 + p.stage0FullY104 + p.stage0FullY112)",
    "p.stage0Y88Depth",
    "p.stage0Y88Width",
    "p.stage0Y96Depth",
    "((double) p.stage0Y96LargestCluster / (p.stage0FullY96 + EPSILON))",
    "((double) p.stage0FullY104 / (p.stage0FullY96 + EPSILON))",
    "((double) p.stage0FullY96 / (p.stage0FullY88 + EPSILON))",
    "p.stage0Y96LargestCluster",
    "p.stage0FullY112",
    "(p.stage0Y88LargestCluster + p.stage0Y96LargestCluster)",
    "p.stage0FullY104",
    "p.stage0Y88LargestCluster",
]

NUM = r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?"


def parse_tree(text, pos, nodes):
    """Parse `if (EXPR <= NUM) { ... } else { ... }` or `return NUM;` starting at pos."""
    ws = re.compile(r"\s*")
    pos = ws.match(text, pos).end()
    m = re.compile(r"return\s+(" + NUM + r")\s*;").match(text, pos)
    if m:
        nodes.append([-1, float(m.group(1)), -1, -1])
        return len(nodes) - 1, m.end()
    m = re.compile(r"if \((.*) <= (" + NUM + r")\) \{").match(text, pos)
    if not m:
        raise ValueError("unexpected token at %d: %r" % (pos, text[pos:pos + 80]))
    expr, thr = m.group(1).strip(), float(m.group(2))
    if expr not in FEATURES:
        raise ValueError("unknown feature expression %r" % expr)
    me = len(nodes)
    nodes.append([FEATURES.index(expr), thr, -1, -1])
    left, pos = parse_tree(text, m.end(), nodes)
    pos = re.compile(r"\s*\}\s*else\s*\{").match(text, pos).end()
    right, pos = parse_tree(text, pos, nodes)
    pos = re.compile(r"\s*\}").match(text, pos).end()
    nodes[me][2], nodes[me][3] = left, right
    return me, pos


def main():
    src = open(sys.argv[1]).read()
    nodes, roots = [], []
    for t in range(16):
        m = re.search(r"inline double tree%d\(int scout, const MonsterFeatures& p\) \{" % t, src)
        root, _ = parse_tree(src, m.end(), nodes)
        roots.append(root)
    out = sys.stdout
    out.write("// Generated by gen_p19_trees.py from native/src/p19_monster_gate.hpp (P19 monster gate,\n")
    out.write("// 16 trees; score = (tree0 + ... + tree15) / 16).  Do not edit by hand.\n")
    out.write("// feature == -1 marks a leaf whose value is the tree output; otherwise go left when\n")
    out.write("// feature_value <= value.  Feature ids: see bsf_p19_feature().\n")
    out.write("struct P19Node { int feature; double value; int left; int right; };\n")
    out.write("static const P19Node kP19Nodes[%d] = {\n" % len(nodes))
    for f, v, l, r in nodes:
        out.write("    {%d, %s, %d, %d},\n" % (f, v.hex(), l, r))
    out.write("};\n")
    out.write("static const int kP19TreeRoot[16] = {%s};\n" % ", ".join(str(r) for r in roots))


if __name__ == "__main__":
    main()
