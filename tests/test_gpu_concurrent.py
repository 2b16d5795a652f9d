"""Plain steps of two solves that overlap on the device.

The plain entry points (launcher, launcher3d, hydro_step2d, hydro_step3d) reduce the new
signal speeds into a state kept per stream: two solves enqueued on different streams,
or launched from two host threads, may run at the same time and must each get the Dt
and fields they get alone.  Solve B is given 40x the energy of solve A, so any mixing of
the two reductions would change A's Dt.
"""
import threading

import numpy as np
import pytest
import torch

from gpu_helpers import assert_bitwise, random_state_2d, random_state_3d, to_dev, to_host

pytestmark = pytest.mark.gpu

DTDX = 0.02  # keeps every updated state physical (see test_gpu_parity.py)


def _pair(make, sizes, seeds):
    out = []
    for nb, seed, scale in zip(sizes, seeds, (1.0, 40.0)):
        rhoE, mom = make(*nb, seed=seed)
        rhoE[..., 1] *= scale
        out.append((nb, to_dev(rhoE), to_dev(mom)))
    return out


def _outputs(rhoE, mom):
    return torch.full_like(rhoE, -1.0), torch.full_like(mom, -1.0), torch.zeros(1, dtype=torch.float64, device="cuda")


def _step(gpu, dim, nb, r, m, out, stream):
    k = gpu.constants(nb, 0.5 if dim == 2 else 0.4)
    n_r, n_m, dt = out
    if dim == 2:
        gpu.step2d(r, m, n_r, n_m, dt, k["C_min"], k["gamma"], k["gmo"], k["divgamma"], k["K"], DTDX, DTDX,
                   stream=stream)
    else:
        gpu.step3d(r, m, n_r, n_m, dt, k["C_min"], k["gamma"], k["gmo"], k["divgamma"], k["K"], DTDX, DTDX, DTDX,
                   stream=stream)


@pytest.mark.parametrize("dim", [2, 3])
def test_steps_on_two_streams_match_each_alone(gpu, dim):
    if dim == 2:
        solves = _pair(random_state_2d, [(1500, 1000), (1200, 1100)], [11, 12])
    else:
        solves = _pair(random_state_3d, [(120, 100, 90), (100, 110, 80)], [13, 14])
    alone = []
    for nb, r, m in solves:
        out = _outputs(r, m)
        _step(gpu, dim, nb, r, m, out, None)
        alone.append([to_host(t) for t in out])
    assert alone[0][2][0] != alone[1][2][0]

    cur = torch.cuda.current_stream()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [_outputs(r, m) for _, r, m in solves]
    for s in streams:
        s.wait_stream(cur)
    for _ in range(30):  # the same step again and again, alternating, so the two overlap
        for (nb, r, m), out, s in zip(solves, outs, streams):
            _step(gpu, dim, nb, r, m, out, s)
    torch.cuda.synchronize()
    for name, want, out in zip("AB", alone, outs):
        got = [to_host(t) for t in out]
        assert_bitwise(got[2], want[2], f"{dim}D solve {name}: Dt_next")
        assert_bitwise(got[0], want[0], f"{dim}D solve {name}: rhoE_next")
        assert_bitwise(got[1], want[1], f"{dim}D solve {name}: momentum_next")


def test_launcher_from_two_threads_matches_each_alone(gpu):
    solves = _pair(random_state_2d, [(1500, 1000), (1200, 1100)], [21, 22])
    consts = [gpu.constants(nb, 0.5) for nb, _, _ in solves]

    def launch(i, out, stream):
        (nb, r, m), k = solves[i], consts[i]
        n_r, n_m, dt = out
        gpu.launcher(r, m, n_r, n_m, dt, k["C_min"], k["gamma"], k["gmo"], k["divgamma"], k["K"], DTDX, DTDX,
                     stream=stream)

    alone = []
    for i, (_, r, m) in enumerate(solves):
        out = _outputs(r, m)
        launch(i, out, None)
        alone.append([to_host(t) for t in out])

    outs = [_outputs(r, m) for _, r, m in solves]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    errors = []

    def worker(i):
        try:
            for _ in range(20):
                launch(i, outs[i], streams[i])
        except Exception as e:  # noqa: BLE001  (reported by the main thread)
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    torch.cuda.synchronize()
    for name, want, out in zip("AB", alone, outs):
        got = [to_host(t) for t in out]
        assert_bitwise(got[2], want[2], f"launcher solve {name}: Dt_next")
        assert_bitwise(got[0], want[0], f"launcher solve {name}: rhoE_next")


def test_more_streams_than_the_state_pool(gpu):
    """300 distinct streams, more than the 256 pooled reduction states of a device: the
    later streams get allocated states, and every stream's step gives the Dt and fields
    of the step alone."""
    import ctypes

    # The HIP runtime calls are resolved through the solver library's own handle, so they
    # reach the runtime it (and torch) use, never a second copy.
    hip = gpu.lib()
    hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    nb = (64, 48)
    rhoE, uv = random_state_2d(*nb, seed=31)
    r, m = to_dev(rhoE), to_dev(uv)
    out = _outputs(r, m)
    _step(gpu, 2, nb, r, m, out, None)
    want = [to_host(t) for t in out]
    streams = []
    try:
        for _ in range(300):
            h = ctypes.c_void_p()
            assert hip.hipStreamCreateWithFlags(ctypes.byref(h), 1) == 0  # hipStreamNonBlocking
            streams.append(h)
        torch.cuda.synchronize()
        for i, h in enumerate(streams):
            got_out = _outputs(r, m)
            torch.cuda.synchronize()
            _step(gpu, 2, nb, r, m, got_out, h.value)
            assert hip.hipStreamSynchronize(h.value) == 0
            got = [to_host(t) for t in got_out]
            assert_bitwise(got[2], want[2], f"stream {i}: Dt_next")
            assert_bitwise(got[0], want[0], f"stream {i}: rhoE_next")
    finally:
        torch.cuda.synchronize()
        for h in streams:
            hip.hipStreamDestroy(h.value)


def test_stream_states_are_released(gpu):
    """More streams than the pool of 256 reduction states: each steps once (the streams
    past the pool get allocated states), then hydro_stream_release gives every state
    back -- pooled ones to the pool, allocated ones freed -- and a stream created
    afterwards steps with a recycled state and gets the same Dt as the others."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    nb = (20, 30)
    rhoE, mom = random_state_2d(*nb, seed=21)
    r, m = to_dev(rhoE), to_dev(mom)
    out = _outputs(r, m)
    _step(gpu, 2, nb, r, m, out, None)
    want = to_host(out[2])[0]
    streams = []
    for _ in range(300):
        h = C.c_void_p()
        assert hip.hipStreamCreate(C.byref(h)) == 0
        streams.append(h.value)
        # A new stream may get the handle of one destroyed earlier in the process whose
        # state is still bound (the case the release call is for): unbind it first.
        gpu.stream_release(h.value)
    bound0, free0 = gpu.stream_states()
    outs = []
    for h in streams:
        outs.append(_outputs(r, m))
        _step(gpu, 2, nb, r, m, outs[-1], h)
    torch.cuda.synchronize()
    bound1, free1 = gpu.stream_states()
    assert bound1 == bound0 + 300 and free1 == max(0, free0 - 300), (bound0, free0, bound1, free1)
    assert all(to_host(o[2])[0] == want for o in outs)
    for h in streams:
        gpu.stream_release(h)
        assert hip.hipStreamDestroy(C.c_void_p(h)) == 0
    assert gpu.stream_states() == (bound0, free0)
    h = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(h)) == 0
    o = _outputs(r, m)
    _step(gpu, 2, nb, r, m, o, h.value)
    torch.cuda.synchronize()
    assert to_host(o[2])[0] == want
    gpu.stream_release(h.value)
    assert hip.hipStreamDestroy(h) == 0
    assert gpu.stream_states() == (bound0, free0)
