"""Extract the reference's recorded compiler statistics into a small fixture (run in the build
container, where the reference tree is mounted; the GPU box never runs this).

Source: results/raw/phase5/sv_*.json and results/raw/explore/*.json of the reference -- one record
per (family, n, depth, P, L, k, n_run, fuse, tiling, layout, build) with stats.{n_ops, n_passes,
n_remaps, n_segments, fused_1q}.  The build's MAXH is builds/<build>/maxh.txt (absent -> 2,
host/sv/runtime.py:86-87).  Output: tests/golden/sched_stats.json (data only).
"""
import glob
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sched_stats.json")


def maxh(build):
    f = os.path.join(REF, "builds", build, "maxh.txt")
    return int(open(f).read()) if os.path.exists(f) else 2


def main():
    seen = {}
    files = sorted(glob.glob(os.path.join(REF, "results/raw/phase5/sv_*.json")))
    files += sorted(glob.glob(os.path.join(REF, "results/raw/explore/*.json")))
    for f in files:
        try:
            d = json.load(open(f))
        except ValueError:
            continue
        recs = d.get("records") if isinstance(d, dict) else None
        for r in recs or []:
            if not isinstance(r, dict) or "stats" not in r or "n_run" not in r:
                continue
            key = (r["family"], r["n"], r.get("depth"), r["P"], r["L"], r["k"], r["n_run"],
                   r.get("fuse", 1), r.get("tiling", 1), r.get("layout", "block"), r.get("xclbin"))
            st = r["stats"]
            val = dict(n_ops=st["n_ops"], n_passes=st["n_passes"], n_remaps=st["n_remaps"],
                       n_segments=st["n_segments"], fused_1q=st["fused_1q"])
            if key in seen:
                assert seen[key][1] == val, ("conflicting records", key)
                continue
            seen[key] = (os.path.relpath(f, REF), val)
    out = []
    for key, (src, val) in sorted(seen.items(), key=lambda kv: str(kv[0])):
        fam, n, depth, P, L, k, n_run, fuse, tiling, layout, build = key
        out.append(dict(family=fam, n=n, depth=depth, P=P, L=L, k=k, n_run=n_run, fuse=int(fuse),
                        tiling=int(tiling), layout=layout, build=build, maxh=maxh(build),
                        source=src, stats=val))
    json.dump(out, open(OUT, "w"), indent=0)
    print(len(out), "records ->", OUT)


if __name__ == "__main__":
    main()
