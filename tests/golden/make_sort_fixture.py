"""Writes sort_ties.json: the permutation libstdc++'s std::sort produces when
sorting records by an integer key with many ties (the reference orders its
train set this way, data/dataset.cpp:17-21, and the order of equal keys fixes
the query-document numbering and summation order).

Run: python tests/golden/make_sort_fixture.py   (needs g++)
"""
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r"""
#include <algorithm>
#include <cstdio>
#include <vector>
struct R { int key; int pos; };
int main() {
    int n; if (std::scanf("%d", &n) != 1) return 1;
    std::vector<R> v(n);
    for (int i = 0; i < n; i++) { std::scanf("%d", &v[i].key); v[i].pos = i; }
    std::sort(v.begin(), v.end(), [](const R& a, const R& b) { return a.key < b.key; });
    for (int i = 0; i < n; i++) std::printf("%d\n", v[i].pos);
}
"""

def main():
    rng = np.random.default_rng(2024)
    keys = rng.integers(0, 37, 3000).tolist()          # many ties, > 16 elements per run
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "s.cpp")
        exe = os.path.join(d, "s")
        open(src, "w").write(SRC)
        subprocess.run(["g++", "-O2", "-o", exe, src], check=True)
        out = subprocess.run([exe], input=f"{len(keys)}\n" + "\n".join(map(str, keys)) + "\n",
                             capture_output=True, text=True, check=True).stdout
    perm = [int(x) for x in out.split()]
    stable = sorted(range(len(keys)), key=lambda i: keys[i])
    json.dump({"keys": keys, "perm": perm, "differs_from_stable": perm != stable},
              open(os.path.join(HERE, "sort_ties.json"), "w"))

if __name__ == "__main__":
    main()
