"""rtkhip_cg_step_noalloc against rtkhip_cg_step through the C ABI: the same CG step
on identical copies of (X, R, P, AP, state) gives bit-identical vectors and FP64
scalars, on two calls in a row, the second on a smaller vector (the shared partial-sum
array keeps the first call's contents, which the second must not read)."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def _vectors(n, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    x, r, p, ap = (torch.randn(n, generator=g, dtype=torch.float32).cuda() for _ in range(4))
    return x, r, p, ap


def _step(fn, n, x, r, p, ap):
    from rtk_amd import _lib
    x, r, p = x.clone(), r.clone(), p.clone()
    state = torch.zeros(4, dtype=torch.float64, device="cuda")
    state[0] = float((r.double() * r.double()).sum())  # ||R||^2 on entry
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    _lib.check(fn(n, ptr(x), ptr(r), ptr(p), ptr(ap), ptr(state), None))
    torch.cuda.synchronize()
    return x, r, p, state


def test_noalloc_cg_step_matches_the_allocating_step(hip):
    from rtk_amd import _lib
    L = _lib.lib()
    # 2^20 + 7 elements take the full 1024-workgroup grid; 3000 take 12 workgroups,
    # so the second call's finalize reads only partial sums that call wrote
    for n, seed in ((2 ** 20 + 7, 1), (3000, 2)):
        print(f"cg_step_noalloc n={n}", flush=True)
        vin = _vectors(n, seed)
        ref = _step(L.rtkhip_cg_step, n, *vin)
        out = _step(L.rtkhip_cg_step_noalloc, n, *vin)
        for a, b, name in zip(out, ref, ("X", "R", "P", "state")):
            assert torch.equal(a, b), (n, name, float((a.double() - b.double()).abs().max()))
        # the step did something: alpha and beta are finite and X moved
        assert torch.isfinite(out[3]).all() and not torch.equal(out[0], vin[0])
