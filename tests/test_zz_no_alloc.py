"""The allocation-free entry point ovl_observables_noalloc (static per-device scratch, device inputs
and outputs, no context) against ovl_system_observables of a context holding the same fields: two
calls in a row on one device, the second on a smaller lattice and another colour group, so that
anything the first call leaves in the shared scratch would show."""
import ctypes as C

import numpy as np
import pytest

from overlap_amd import configs as cf
from overlap_amd import ovl

pytestmark = pytest.mark.gpu


def _fields(nc, n, seed):
    g = cf.rng(seed)
    links = cf.su2_links(n, 0.6, g) if nc == 2 else cf.su3_links(n, 0.6, g)
    e = cf.efield(n, nc, 0.5, g)
    # links pushed off the group by their own factors 1 + O(1e-2), so the unitarity entries are not ~0
    per = 4 if nc == 2 else 18
    lam = 1.0 + g.normal(0, 0.01, size=3 * int(np.prod(n)))
    return (links.reshape(-1, per) * lam[:, None]).reshape(-1), e


class _Hip:
    """The HIP runtime the library itself loaded (same soname, so dlopen returns that handle): the
    test's own device buffers for the inputs and outputs."""

    def __init__(self):
        ovl.lib()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def put(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a, np.float64)
        p = C.c_void_p()
        self._ok(self.rt.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)), "hipMalloc")
        self.bufs.append(p)
        self._ok(self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1), "hipMemcpy H2D")
        return p.value

    def get(self, ptr: int, n: int) -> np.ndarray:
        out = np.empty(n)
        self._ok(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self._ok(self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2),
                 "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def test_noalloc_observables_match_the_context_entry_point(gpu):
    hip = _Hip()
    nobs = len(ovl.OBS_NAMES)
    obs_dev = hip.put(np.full(nobs, np.nan))
    for nc, n, seed in [(3, (6, 5, 7), 71), (2, (4, 3, 5), 72)]:
        print(f"[noalloc] nc={nc} n={n}", flush=True)
        links, e = _fields(nc, n, seed)
        with ovl.Context(nc, n, 16) as ctx:
            ctx.set_links(links)
            ctx.set_efield(e)
            want = np.array([ctx.system_observables(0)[k] for k in ovl.OBS_NAMES])
        ld, ed = hip.put(links), hip.put(e)
        ovl.observables_noalloc(nc, n, ld, ed, obs_dev, 0)
        got = hip.get(obs_dev, nobs)
        assert np.isfinite(got).all()
        assert abs(got[4]) > 1e-3  # unitarity violation of the scaled links
        # same kernels, same fixed reduction grid: the sums agree bit for bit; the two rms values
        # and dN_CS/dt are one IEEE sqrt / division away from them, on the device here
        direct = [0, 1, 2, 4]
        assert np.array_equal(got[direct], want[direct]), (got, want)
        np.testing.assert_allclose(got, want, rtol=1e-15, atol=0)
    hip.free()
