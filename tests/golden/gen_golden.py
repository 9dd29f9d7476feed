"""Write the golden count maps (tests/golden/golden.npz + golden.json).

Counts come from MDZ's own formula loops (src/frac_*.c compiled as they are
into oracle/_ref/libfrac_ref.so by oracle/Makefile), driven per pixel by the
oracle's restatement of the line workers (src/fractal.c:29-257).  Run in a
container that has the reference tree:

    make -C oracle all && python tests/golden/gen_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests import cases  # noqa: E402
from mdz_amd import workloads as WL  # noqa: E402  (pure-Python view descriptions)

OUT = os.path.dirname(os.path.abspath(__file__))


def vdict(rect, precision, depth, W, H, fractal=0, family=0, path=0, julia=(None, None)):
    return dict(rect=list(rect), precision=precision, depth=depth, width=W, height=H,
                fractal=fractal, family=family, path=path, julia=list(julia))


def cases_list():
    out = []
    # config 1 (double and long double paths), every 4th row
    v1 = WL.config1_view()
    r1 = [s.decode() for s in v1._keep[:4]]
    for path in (O.PATH_DOUBLE, O.PATH_LD):
        for fr in range(4):
            out.append((f"config1_{'double' if path == 2 else 'ld'}_f{fr}",
                        vdict(r1, 64, 1000, 240, 180, fractal=fr, path=path),
                        (0, 45, 4), (0, 240)))
    # gallery scenes
    for name in cases.SCENES:
        rect, p, depth, fr, path, (W, H) = cases.scene_rect(name)
        ch, cw = min(24, H), 32
        out.append((f"scene_{name}", vdict(rect, p, depth, W, H, fractal=fr, path=path),
                    (H // 2 - ch // 2, ch, 1), (W // 2 - cw // 2, cw)))
    # other formulas on two scenes
    for name in ("subdudehue", "test"):
        rect, p, depth, _, path, (W, H) = cases.scene_rect(name)
        for fr in (1, 2, 3):
            out.append((f"scene_{name}_f{fr}", vdict(rect, p, depth, W, H, fractal=fr),
                        (60, 16, 1), (100, 24)))
    # Julia family
    J = cases.JULIA
    for fr in range(4):
        out.append((f"julia_f{fr}", vdict(J["rect"], J["precision"], J["depth"], 96, 72,
                                          fractal=fr, family=1, julia=J["c"]),
                    (0, 24, 3), (0, 96)))
    # config 2 and config 5 crops
    v2 = WL.config2_view()
    out.append(("config2_crop", vdict([s.decode() for s in v2._keep[:4]], 128, 5000, 640, 480),
                (224, 8, 1), (300, 16)))
    v5 = WL.config5_view()
    out.append(("config5_crop", vdict([s.decode() for s in v5._keep[:4]], 1024, 20000, 3840, 2160),
                (1076, 4, 1), (1916, 8)))
    # config 3: crops of the first tiles with depth <= 4000
    n = 0
    for sp in WL.config3_specs():
        if sp["depth"] > 4000 or sp["precision"] == 53:
            continue
        v = WL.spec_view(sp)
        out.append((f"config3_tile{sp['index']}",
                    vdict([s.decode() for s in v._keep[:4]], sp["precision"], sp["depth"],
                          sp["side"], sp["side"], fractal=sp["fractal"]),
                    (sp["side"] // 2 - 4, 8, 1), (sp["side"] // 2 - 4, 8)))
        n += 1
        if n == 4:
            break
    return out


def main():
    O.build()
    O.use_reference(True)          # MDZ's own loops
    arrays, meta = {}, []
    for name, vd, lines, cols in cases_list():
        counts = O.render(vd["rect"], precision=vd["precision"], depth=vd["depth"],
                          width=vd["width"], height=vd["height"], fractal=vd["fractal"],
                          family=vd["family"], path=vd["path"],
                          julia=tuple(vd["julia"]), lines=lines, cols=cols,
                          threads=os.cpu_count() or 4)
        arrays[name] = counts.astype(np.int32)
        meta.append(dict(name=name, view=vd, lines=list(lines), cols=list(cols)))
        print(f"{name}: {counts.shape} distinct={len(np.unique(counts))} "
              f"max={counts.max()} zeros={(counts == 0).mean():.2f}", flush=True)
    np.savez_compressed(os.path.join(OUT, "golden.npz"), **arrays)
    with open(os.path.join(OUT, "golden.json"), "w") as f:
        json.dump(dict(source="src/frac_*.c compiled as-is (oracle/_ref), -O0 -ggdb -std=gnu99",
                       cases=meta), f, indent=1)


if __name__ == "__main__":
    main()
