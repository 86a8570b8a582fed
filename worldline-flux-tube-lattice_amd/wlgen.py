#!/usr/bin/env python3
"""Loop-ensemble generator: writes seeded unit loops to a file.

    python3 worldline-flux-tube-lattice_amd/wlgen.py --config-id 1 --Nl 5120 --N 1024 --format text
        -> worldlines.5120.1024.2.dat, the text layout getwl.c:18-28 reads and
           the file name EffAct.c:314 opens
    python3 worldline-flux-tube-lattice_amd/wlgen.py --config-id 2 --Nl 5120 --N 4096 --format bin
        -> worldlines.5120.4096.2.bin, SoA float32 with a header and a content
           hash (wfl_loops_read verifies both)

The loops come from the host generator (Philox4x32-10 keyed by config id,
Box-Muller increments of variance 2/N per component, closed, centre of mass
at the origin), the same ensemble wfl_loops_generate puts on the device.
--verify reads the file back and checks it against the generated arrays.
"""
import argparse
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _wfl():
    spec = importlib.util.spec_from_file_location("wfl", os.path.join(HERE, "wfl.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules.setdefault("wfl", mod)
    spec.loader.exec_module(mod)
    return mod


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--config-id", type=int, required=True, help="generator seed (key 0x5EED0000 + id)")
    ap.add_argument("--Nl", type=int, required=True, help="number of loops")
    ap.add_argument("--N", type=int, required=True, help="points per loop (a multiple of 4)")
    ap.add_argument("--format", choices=("text", "bin"), default="bin")
    ap.add_argument("--out", help="output path (default: worldlines.<Nl>.<N>.2.dat or .bin)")
    ap.add_argument("--verify", action="store_true", help="read the file back and compare")
    a = ap.parse_args(argv)
    wfl = _wfl()
    out = a.out or f"worldlines.{a.Nl}.{a.N}.2.{'dat' if a.format == 'text' else 'bin'}"
    x, y = wfl.gen_loops(a.config_id, a.Nl, a.N)
    if a.format == "text":
        wfl.write_text_loops(out, x, y)
    else:
        wfl.write_loops(out, x, y, config_id=a.config_id)
    if a.verify:
        if a.format == "text":
            rx, ry = wfl.read_text_loops(out, a.Nl, a.N)
        else:
            rx, ry, cid = wfl.read_loops(out)
            if cid != a.config_id:
                raise SystemExit(f"{out}: config id {cid}, expected {a.config_id}")
        if not (np.array_equal(rx, x) and np.array_equal(ry, y)):
            raise SystemExit(f"{out}: read-back differs from the generated loops")
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
