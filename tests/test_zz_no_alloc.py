"""tf_time_integration_noalloc, the loop without a workspace argument, through
the C ABI: on two calls in a row (the second on a smaller grid) it gives the
results of tf_time_integration with a workspace bit for bit (state, step
count, t, flags), it leaves the device's free memory unchanged once its
array is loaded, and it rejects a grid beyond the supported size before any
launch."""

import ctypes

import pytest
import torch

from tinyfluids_amd import ics

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def loop(name, state, scratch, dx, gamma, t_final, max_steps, hb, ws=None):
    import tinyfluids_amd._lib as L
    _, nx, ny, nz = state.shape
    n, t, in_s, nf = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
    args = [state.data_ptr(), scratch.data_ptr(), nx, ny, nz, dx, t_final, gamma, int(hb), max_steps, 0.4]
    if ws is not None:
        args.append(ws.data_ptr())
    args += [ctypes.byref(n), ctypes.byref(t), ctypes.byref(in_s), ctypes.byref(nf),
             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)]
    L.check(getattr(L.load(), name)(*args))
    return (scratch if in_s.value else state), n.value, t.value, in_s.value, nf.value


CASES = [  # (shape, t_final, max_steps, open boundaries): the second call is the smaller one
    ((40, 36, 96), 0.02, -1, False),
    ((12, 20, 16), float("inf"), 9, True),
]


def test_noalloc_loop_equals_workspace_loop():
    import tinyfluids_amd._lib as L
    lib = L.load()
    for k, (shape, t_final, max_steps, hb) in enumerate(CASES):
        print(f"noalloc case {k}: {shape} t_final={t_final} max_steps={max_steps} open={hb}", flush=True)
        s0, dx = ics.blast(shape, centre=(0.45, 0.5, 0.55), radius=0.15, p_in=20.0)
        x = torch.from_numpy(s0).to(DEV)
        ws = torch.empty(lib.tf_workspace_bytes(*shape), dtype=torch.uint8, device=DEV)
        ref_bufs = (x.clone(), torch.empty_like(x))
        ref = loop("tf_time_integration", *ref_bufs, dx, ics.GAMMA_BLAST, t_final, max_steps, hb, ws)
        got_bufs = [(x.clone(), torch.empty_like(x)) for _ in range(3)]
        torch.cuda.synchronize()
        got = loop("tf_time_integration_noalloc", *got_bufs[0], dx, ics.GAMMA_BLAST, t_final, max_steps, hb)
        torch.cuda.synchronize()
        # once the array is loaded, further calls leave the free memory as it is
        free0 = torch.cuda.mem_get_info()[0]
        again = [loop("tf_time_integration_noalloc", *b, dx, ics.GAMMA_BLAST, t_final, max_steps, hb)
                 for b in got_bufs[1:]]
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0
        for g in [got] + again:
            assert g[1:] == ref[1:], (g[1:], ref[1:])
            assert g[1] > 0 and not g[4]
            assert torch.equal(g[0], ref[0])


def test_noalloc_rejects_an_unsupported_grid():
    import tinyfluids_amd._lib as L
    x = torch.zeros(5, 2, 2, 2, device=DEV)
    n, t, in_s, nf = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
    rc = L.load().tf_time_integration_noalloc(x.data_ptr(), x.data_ptr() + 4, 1025, 1024, 1024, 1e-3, 1.0, 1.4, 0, 1,
                                              0.4, ctypes.byref(n), ctypes.byref(t), ctypes.byref(in_s),
                                              ctypes.byref(nf), None)
    assert rc == L.TF_ERR_UNSUPPORTED
    assert b"2^30" in L.load().tf_last_error()
