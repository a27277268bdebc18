"""The allocation-free waveform measurement entry point
(fwat_waveform_adjoint_batch_noalloc_) against fwat_waveform_adjoint_batch_,
through the C ABI: two calls in a row on the same GPU, the second with fewer
traces, shorter traces and fewer windows, so results left in the shared
scratch by the first call must not leak into the second.  The caller's device
buffers come from the HIP runtime the library itself links (the caller owns
them; torch would bring a second runtime into the process); the entry point
allocates none."""
import ctypes as C

import numpy as np
import pytest

from specfwat_amd import _lib
from specfwat_amd.misfit import bandpass_sos

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Caller-owned device memory through libamdhip64 (hipMalloc / hipMemcpy / hipFree)."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            _lib.lib()  # the library's own runtime is the one already loaded
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), self.host.nbytes) == 0
        assert self.hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0  # host -> device

    @property
    def address(self):
        return self.ptr.value

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device -> host
        return out

    def free(self):
        if self.ptr:
            assert self.hip().hipFree(self.ptr) == 0
            self.ptr = C.c_void_p()


def case(rng, nstep, ntrace, windows, dt=0.01):
    t = np.arange(nstep) * dt
    obs = np.zeros((nstep, ntrace), np.float32)
    syn = np.zeros((nstep, ntrace), np.float32)
    for k in range(ntrace):
        f = rng.uniform(0.5, 2.0)
        tc = rng.uniform(0.3, 0.7) * nstep * dt
        obs[:, k] = np.sin(2 * np.pi * f * t) * np.exp(-((t - tc) / 2.0) ** 2)
        syn[:, k] = 1.1 * np.sin(2 * np.pi * f * (t - 0.05)) * np.exp(-((t - tc) / 2.0) ** 2)
    return obs, syn, np.asarray(windows, np.float64)


def reference(nstep, ntrace, dt, obs, syn, win, sos):
    mis = np.zeros(ntrace)
    res = np.zeros((ntrace, len(win)))
    adj = np.zeros_like(obs)
    _lib.call("fwat_waveform_adjoint_batch_", nstep, ntrace, dt, obs, syn, len(win), win, 0.05, 1,
              sos.shape[0], sos if sos.size else np.zeros(6, np.float32), mis, res, adj)
    return mis, res, adj


def no_alloc(nstep, ntrace, dt, obs, syn, win, sos, with_residuals=True):
    d_obs, d_syn = DeviceBuffer(obs), DeviceBuffer(syn)
    d_adj = DeviceBuffer(np.full((nstep, ntrace), np.nan, np.float32))
    d_mis = DeviceBuffer(np.full(ntrace, np.nan))
    d_res = DeviceBuffer(np.full((ntrace, len(win)), np.nan)) if with_residuals else None
    _lib.call("fwat_waveform_adjoint_batch_noalloc_", nstep, ntrace, dt, d_obs.address, d_syn.address,
              len(win), win, 0.05, 1, sos.shape[0], sos if sos.size else np.zeros(6, np.float32),
              d_mis.address, None if d_res is None else d_res.address, d_adj.address, 0)
    out = (d_mis.get(), None if d_res is None else d_res.get(), d_adj.get())
    for b in (d_obs, d_syn, d_adj, d_mis, d_res):
        if b is not None:
            b.free()
    return out


def test_no_alloc_waveform_matches_allocating_entry_point():
    rng = np.random.default_rng(17)
    dt = 0.01
    calls = [
        # (nstep, ntrace, windows, filtered, residuals)
        (3000, 700, [[2.0, 12.0], [9.0, 25.0], [20.0, 29.0]], True, True),
        (1200, 90, [[1.0, 10.0]], False, False),
        (1500, 64, [[0.5, 7.5], [4.0, 14.0]], True, True),
    ]
    for nstep, ntrace, windows, filtered, with_res in calls:
        print(f"no-alloc call: nstep {nstep}, ntrace {ntrace}, {len(windows)} windows, filter {filtered}",
              flush=True)
        obs, syn, win = case(rng, nstep, ntrace, windows, dt)
        sos = bandpass_sos(dt, 0.1, 2.5) if filtered else np.zeros((0, 6), np.float32)
        m0, r0, a0 = reference(nstep, ntrace, dt, obs, syn, win, sos)
        m1, r1, a1 = no_alloc(nstep, ntrace, dt, obs, syn, win, sos, with_res)
        assert np.all(np.isfinite(m1)) and np.all(np.isfinite(a1))
        em = np.max(np.abs(m1 - m0) / np.maximum(np.abs(m0), 1e-300))
        ea = np.max(np.abs(a1 - a0)) / np.max(np.abs(a0))
        print(f"  misfits rel {em:.2e}, adjoint sources {ea:.2e} of max, bitwise "
              f"{np.array_equal(m1, m0) and np.array_equal(a1, a0)}")
        assert em < 1e-12 and ea < 1e-6
        if with_res:
            assert np.max(np.abs(r1 - r0)) <= 1e-12 * np.max(np.abs(r0))


def test_no_alloc_refuses_sizes_beyond_its_scratch():
    """Sizes past the scratch limits are refused on the host before any device work
    (the library exits with a message; run in a child process)."""
    import subprocess
    import sys
    code = ("import numpy as np\n"
            "from specfwat_amd import _lib\n"
            "_lib.call('fwat_waveform_adjoint_batch_noalloc_', 100, 20000, 0.01, 1, 1, 1,"
            " np.array([0.1, 0.5]), 0.05, 1, 0, np.zeros(6, np.float32), 1, None, 1, 0)\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       cwd=__import__("os").path.dirname(__import__("os").path.dirname(__file__)))
    assert r.returncode != 0 and "at most 16384" in r.stderr, r.stderr
