"""Shipped tile-shape table of the integer search, keyed by (PU shape, search range).

The autotuner (``xme_tune``) times the candidate tile shapes (K rows per lane,
LC lanes per candidate row) on the device; its winners for the shapes and
ranges the encoder uses are kept in ``tuned_tiles.json`` beside this module so
that a process installs them with ``xme_tune_set`` instead of re-timing.
Shapes or ranges missing from the table are tuned on first use.

Regenerate on an MI355X with::

    python -m xavs2_amd.tiles --range 64 --range 32 --out xavs2_amd/tuned_tiles.json
"""
import argparse
import json
import os

from . import lib as X

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_tiles.json")


def load(path=TABLE_PATH):
    """{(part, R): (K, LC)} from a table file (empty when absent)."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        doc = json.load(f)
    return {(int(e["part"]), int(e["range"])): (int(e["K"]), int(e["LC"])) for e in doc["tiles"]}


def install(params, parts, stream=None, retune=False, path=TABLE_PATH):
    """Installs the shipped choices for `parts` at params.search_range and tunes
    the rest on this device.  Returns ({part: (K, LC)}, list of parts tuned)."""
    R = params.search_range
    table = {} if retune else load(path)
    missing = []
    for p in parts:
        if (int(p), R) in table:
            X.tune_set(p, R, *table[(int(p), R)])
        else:
            missing.append(int(p))
    if missing:
        X.tune(params, missing, stream)
    return {int(p): X.tune_get(p, R) for p in parts}, missing


def main():
    ap = argparse.ArgumentParser(description="Tune the integer-search tiles on this GPU and write the table.")
    ap.add_argument("--range", type=int, action="append", required=True)
    ap.add_argument("--out", default=TABLE_PATH)
    args = ap.parse_args()
    entries = []
    s = X.Stream()
    for R in args.range:
        prm = X.me_params(1920, 1080, R, 4 * 65536)
        got = X.tune(prm, list(range(25)), s.h)
        for p, (k, lc) in sorted(got.items()):
            entries.append({"part": p, "w": X.PU_W[p], "h": X.PU_H[p], "range": R, "K": k, "LC": lc})
    with open(args.out, "w") as f:
        json.dump({"device": "MI355X (gfx950)", "tiles": entries}, f, indent=1)
    print(f"wrote {len(entries)} entries to {args.out}")


if __name__ == "__main__":
    main()
