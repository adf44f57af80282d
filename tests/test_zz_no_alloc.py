"""The no-allocation entry point tcme_global_mv_static: the same results as
tcme_global_mv (which grows a workspace in the context) on two calls in a
row, the second smaller, both checked against the oracle; a call whose
coarse-SAD table would not fit the fixed scratch array is refused."""
import ctypes as C

import numpy as np
import pytest

from helpers import moving_pair

pytestmark = pytest.mark.gpu


def _run(eng, fn, pairs):
    import torch
    from tcodec_amd import me
    p = torch.tensor(np.asarray(pairs, np.int32).reshape(-1, 2)).cuda()
    out = torch.full((p.shape[0], 3), -99, dtype=torch.int32, device="cuda")
    rc = fn(eng.h, C.c_void_p(p.data_ptr()), p.shape[0], C.c_void_p(out.data_ptr()), None)
    torch.cuda.synchronize()
    return rc, out.cpu().numpy(), me


def test_global_mv_static_equals_allocating_entry_point(orc):
    from tcodec_amd import me, synth
    print("no-alloc global mv", flush=True)
    eng = me.Engine(0)
    try:
        rng = np.random.default_rng(17)
        planes, pairs, want = [], [], []
        for (h, w) in [(240, 320)] * 4 + [(360, 640)] * 2 + [(61, 97)]:
            shift = tuple(int(v) for v in rng.integers(-40, 41, 2))
            ref, cur = moving_pair(h, w, int(rng.integers(0, 1000)), shift=shift)
            rp, cp = eng.pic(ref), eng.pic(cur)
            planes += [rp, cp]
            pairs.append((cp, rp))
            want.append(orc.global_mv(cur, ref))
        sc = synth.Scene(1920, 1080, seed=3, rects=16, pan=(8.0, -4.0))
        f0, f2 = sc.frame(0)[0], sc.frame(2)[0]
        rp, cp = eng.pic(f0), eng.pic(f2)
        planes += [rp, cp]
        pairs.append((cp, rp))
        want.append(orc.global_mv(f2, f0))
        L = eng.L
        # first call: every pair; second call: a smaller input (2 pairs)
        for sel in (list(range(len(pairs))), [5, 1]):
            sub = [pairs[i] for i in sel]
            rc, a, _ = _run(eng, L.tcme_global_mv, sub)
            assert rc == 0
            rc, b, _ = _run(eng, L.tcme_global_mv_static, sub)
            assert rc == 0, L.tcme_last_error().decode()
            assert np.array_equal(a, b), (sel, a, b)
            assert (b[:, 2] == 0).all()
            assert [tuple(r[:2]) for r in b] == [want[i] for i in sel]
        # 120 pairs of 1920x1080 pictures need ~18 MB of coarse SADs: refused
        rc, c, _ = _run(eng, L.tcme_global_mv_static, [pairs[-1]] * 120)
        assert rc == -1 and "scratch" in L.tcme_last_error().decode()
        assert (c == -99).all()
        # a bad pair still gets its status, as with tcme_global_mv
        rc, d, _ = _run(eng, L.tcme_global_mv_static, [(planes[0], planes[-1]), (9999, planes[0])])
        assert rc == 0 and d[:, 2].tolist() == [-1, -4]
    finally:
        eng.close()
