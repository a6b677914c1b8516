"""What the trained direction network gains over Beamer's rules
(hybrid_bfs.cc:167, 178-179), on a graph it never saw: RMAT scale 21,
generator seed 404 (the training graphs use seeds 101 and 202, the margin
was chosen on seed 303), the first 8 lcg(0) roots.  Per root: the device
time of the search with Beamer's choices, with the network's, and the best
of every single-step flip of Beamer's choices (what a direction oracle that
changes one step could reach).  Directions never change the answer, so the
depths are compared too."""
import json
import os

import numpy as np
import pytest

import beedrill_amd as bd

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "beedrill_amd", "models", "direction_mlp_32x32.f32")


def test_learned_policy_gain_on_heldout_graph(ctx):
    ctx.generate_rmat(21, 16, seed=404)
    ctx.mlp_set(np.fromfile(MODEL, np.float32), 32, 32)
    with open(os.path.splitext(MODEL)[0] + ".json") as f:
        margin = json.load(f)["margin_milli"]
    old = ctx.get_tuning()["learned_margin_milli"]
    ctx.set_tuning(learned_margin_milli=int(margin))
    try:
        roots = [int(r) for r in ctx.pick_roots(8)]
        for r in roots[:2]:
            ctx.bfs(r)

        def t(root, policy):
            return min(ctx.bfs(root, 15, 18, policy)["ms"] for _ in range(3))

        beamer = learned = best = 0.0
        for r in roots:
            b = t(r, "beamer")
            dirs = [x["dir"] for x in ctx.levels()]
            depth_b, _ = ctx.fetch(parent=False)
            lt = t(r, "learned")
            depth_l, _ = ctx.fetch(parent=False)
            np.testing.assert_array_equal(depth_l, depth_b)
            flips = []
            for L in range(len(dirs)):
                sc = list(dirs)
                sc[L] = 1 - sc[L]
                ctx.set_script(np.array(sc, np.int32))
                flips.append(t(r, "script"))
            beamer += b
            learned += lt
            best += min([b, lt] + flips)
        print(f"held-out RMAT s21 seed 404, 8 roots: Beamer {beamer * 1e3:.1f} us, "
              f"learned {learned * 1e3:.1f} us ({beamer / learned:.3f}x), "
              f"best single flip {best * 1e3:.1f} us ({beamer / best:.3f}x)")
        assert best <= min(beamer, learned)
        assert learned <= beamer * 1.02
    finally:
        ctx.set_tuning(learned_margin_milli=old)
