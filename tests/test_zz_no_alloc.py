"""ldpc_decode_llrs_device_noalloc against ldpc_decode_llrs_device_async.

The allocation-free entry point keeps the decode kernels' scratch (queue
counter, tally counters, histogram) in a fixed device array inside libldpc.
It must give the context-based decode's outputs bit for bit, on two calls in
a row with the second one smaller (the array keeps the first call's
contents, so anything not cleared would show), for the specialised AR4JA
r1/2 K=1024 kernel and for the nine-code kernels in wave and group mode.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import channel_params

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    return C.CDLL("libamdhip64.so.7")


class DeviceBuffers:
    def __init__(self, hip, n, n_cw):
        self.hip = hip
        self.ptrs = []
        self.llr_in = self._alloc(2 * n * n_cw)
        self.llr_out = self._alloc(2 * n * n_cw)
        self.iters = self._alloc(4 * n_cw)
        self.ok = self._alloc(n_cw)

    def _alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.ptrs.append(p)
        return p

    def upload(self, llr):
        assert self.hip.hipMemcpy(self.llr_in, llr.ctypes.data_as(C.c_void_p), C.c_size_t(llr.nbytes), H2D) == 0

    def download(self, n_cw, n):
        out = np.empty((n_cw, n), dtype=np.uint16)
        it = np.empty(n_cw, dtype=np.uint32)
        ok = np.empty(n_cw, dtype=np.uint8)
        for host, dev in ((out, self.llr_out), (it, self.iters), (ok, self.ok)):
            assert self.hip.hipMemcpy(host.ctypes.data_as(C.c_void_p), dev, C.c_size_t(host.nbytes), D2H) == 0
        return out, it, ok

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


CASES = [
    (1024, "1/2", 0, 2.0, "specialised r1/2 K=1024 kernel"),
    (1024, "2/3", 0, 2.6, "nine-code kernel, wave mode"),
    (4096, "1/2", 0, 1.5, "nine-code kernel, group mode"),
]


@pytest.mark.parametrize("k_info,rate,variant,ebn0,what", CASES, ids=[c[4] for c in CASES])
def test_noalloc_matches_context_decode(hip, oracle, k_info, rate, variant, ebn0, what):
    from ldpc_mi355x import _lib as L
    from ldpc_mi355x import codes, decoder
    lib = L.load()
    code = codes.ar4ja(k_info, rate)
    desc = decoder.ar4ja_desc(k_info, rate)
    cfg = L.DecoderCfg(L.LDPC_PROFILE_T, 200, 0.9375, 0.0, 0.0, 0.0, 0.0, variant)
    mu, sg = channel_params(ebn0, code.k / code.n_tx)
    n_first = 600 if k_info == 1024 else 96
    llr_all = oracle.gen_channel(code, 0x5EED, 0, n_first, mu, sg)
    bufs_a = DeviceBuffers(hip, code.n, n_first)
    bufs_b = DeviceBuffers(hip, code.n, n_first)
    try:
        with decoder.Engine(k_info, rate, max_iter=200, beta=0.9375) as eng:
            for call, n_cw in enumerate((n_first, n_first // 3 + 1)):
                llr = np.ascontiguousarray(llr_all[:n_cw])
                if call == 1:  # the second call decodes other codewords
                    llr = np.ascontiguousarray(llr_all[n_first - n_cw:])
                print(f"no-alloc decode: {what}, call {call}, {n_cw} codewords", flush=True)
                bufs_a.upload(llr)
                bufs_b.upload(llr)
                rc = lib.ldpc_decode_llrs_device_noalloc(C.byref(desc), C.byref(cfg), 0, bufs_a.llr_in, n_cw,
                                                         bufs_a.llr_out, bufs_a.iters, bufs_a.ok, None)
                assert rc == 0, lib.ldpc_last_error(None)
                assert hip.hipDeviceSynchronize() == 0
                eng.decode_llrs_device_async(bufs_b.llr_in.value, n_cw, bufs_b.llr_out.value, bufs_b.iters.value,
                                             bufs_b.ok.value)
                eng.sync()
                got = bufs_a.download(n_cw, code.n)
                exp = bufs_b.download(n_cw, code.n)
                for g, e in zip(got, exp):
                    np.testing.assert_array_equal(g, e)
                if call == 1:  # and both equal the oracle
                    o_out, o_it, o_ok = oracle.decode_t(code, llr, 200, 0.9375)
                    np.testing.assert_array_equal(got[0], o_out)
                    np.testing.assert_array_equal(got[1], o_it)
                    np.testing.assert_array_equal(got[2], o_ok)
    finally:
        bufs_a.free()
        bufs_b.free()


def test_noalloc_rejects_what_it_cannot_hold():
    from ldpc_mi355x import _lib as L
    from ldpc_mi355x import decoder
    lib = L.load()
    desc = decoder.ar4ja_desc(1024, "1/2")
    dummy = C.c_void_p(16)
    too_many = L.DecoderCfg(L.LDPC_PROFILE_T, L_MAX_ITER + 1, 0.9375)
    assert lib.ldpc_decode_llrs_device_noalloc(C.byref(desc), C.byref(too_many), 0, dummy, 1, dummy, dummy, dummy,
                                               None) == L.LDPC_ERR_ARG
    assert b"max_iter" in lib.ldpc_last_error(None)
    prof_r = L.DecoderCfg(L.LDPC_PROFILE_R, 16, 0.0)
    assert lib.ldpc_decode_llrs_device_noalloc(C.byref(desc), C.byref(prof_r), 0, dummy, 1, dummy, dummy, dummy,
                                               None) == L.LDPC_ERR_UNSUPPORTED
    ok_cfg = L.DecoderCfg(L.LDPC_PROFILE_T, 200, 0.9375)
    assert lib.ldpc_decode_llrs_device_noalloc(C.byref(desc), C.byref(ok_cfg), 0, dummy, 0, dummy, dummy, dummy,
                                               None) == 0  # nothing to do
    assert lib.ldpc_decode_llrs_device_noalloc(C.byref(desc), C.byref(ok_cfg), 0, None, 1, dummy, dummy, dummy,
                                               None) == L.LDPC_ERR_ARG


L_MAX_ITER = 8000  # LDPC_NOALLOC_MAX_ITER of include/ldpc_mi355x.h
