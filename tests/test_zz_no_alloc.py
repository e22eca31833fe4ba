"""The allocation-free event entry point (crysp_run_events_noalloc) against the allocating
one (crysp_run_events) on the same deposits, through the C-ABI: two calls in a row, the
second with a smaller input, so the shared scratch array is reused; the size limit and
host arrays are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from crysplight_amd import _lib, events as E
from crysplight_amd.params import Box, Emission, SipmGrid, readout
from tests import cases as K

pytestmark = pytest.mark.gpu

RO8 = readout(SipmGrid(3.2, 8, 8))


def _noalloc(ctx, dep, xyzt, nph, ev, out, seed, tau, pid0=0):
    lib = _lib.load()
    f = lib.crysp_run_events_noalloc
    maps, tmin, npe = out
    ntot = C.c_int64()
    em = Emission(tau, 0.0)
    _lib.ready(xyzt, nph, ev, maps, tmin, npe)
    rc = f(ctx.handle, C.byref(K.C2_BOX), C.byref(K.C2_OP), C.byref(RO8), xyzt.data_ptr(),
           nph.data_ptr(), ev.data_ptr(), int(xyzt.shape[0]), int(dep.n_events), seed, pid0,
           100_000, C.byref(em), maps.data_ptr(), tmin.data_ptr(), npe.data_ptr(), C.byref(ntot))
    return rc, ntot.value


def _tensors(dep, dev):
    import torch
    inp = tuple(torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                for a in (dep.xyzt, dep.nph, dep.event))
    nev = dep.n_events
    out = (torch.full((nev, 8, 8), -7, dtype=torch.int32, device=dev),
           torch.full((nev, 8, 8), -7.0, dtype=torch.float32, device=dev),
           torch.full((nev,), -7, dtype=torch.int32, device=dev))
    return inp, out


def test_noalloc_equals_allocating_entry_point_twice(ctx):
    import torch
    dev = torch.device("cuda", 0)
    bgo = E.GammaMaterial("xcom_BGO.csv", 7.13)
    big, _ = E.gamma_events(K.C2_BOX, bgo, 8500.0, n_events=300, seed=21)
    small, _ = E.gamma_events(K.C2_BOX, bgo, 8500.0, n_events=40, seed=22)
    for dep, seed in ((big, 21), (small, 22)):
        print(f"noalloc call: {dep.n_events} events, {dep.xyzt.shape[0]} deposits", flush=True)
        ref = E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=seed, tau_ns=K.C2_TAU, ctx=ctx)
        (xyzt, nph, ev), out = _tensors(dep, dev)
        rc, ntot = _noalloc(ctx, dep, xyzt, nph, ev, out, seed, K.C2_TAU)
        assert rc == 0, _lib.load().crysp_last_error().decode()
        assert ntot == dep.n_photons == ref.n_photons
        nev = dep.n_events
        assert np.array_equal(out[0].cpu().numpy().reshape(nev, -1), ref.maps.reshape(nev, -1))
        assert np.array_equal(out[1].cpu().numpy().reshape(nev, -1), ref.tmin.reshape(nev, -1))
        assert np.array_equal(out[2].cpu().numpy(), ref.npe)


def test_noalloc_refuses_oversize_and_host_arrays(ctx):
    import torch
    dev = torch.device("cuda", 0)
    bgo = E.GammaMaterial("xcom_BGO.csv", 7.13)
    dep, _ = E.gamma_events(K.C2_BOX, bgo, 8500.0, n_events=8, seed=3)
    (xyzt, nph, ev), out = _tensors(dep, dev)
    lib = _lib.load()
    f = lib.crysp_run_events_noalloc
    em = Emission(K.C2_TAU, 0.0)
    maps, tmin, npe = out
    # more deposits than the scratch array holds: refused before any launch
    rc = f(ctx.handle, C.byref(K.C2_BOX), C.byref(K.C2_OP), C.byref(RO8), xyzt.data_ptr(),
           nph.data_ptr(), ev.data_ptr(), (1 << 20) + 1, dep.n_events, 1, 0, 100_000,
           C.byref(em), maps.data_ptr(), tmin.data_ptr(), npe.data_ptr(), None)
    assert rc == -4 and "deposits" in lib.crysp_last_error().decode()
    # host arrays would need staging buffers: refused
    h = np.ascontiguousarray(dep.xyzt)
    rc = f(ctx.handle, C.byref(K.C2_BOX), C.byref(K.C2_OP), C.byref(RO8), h.ctypes.data,
           nph.data_ptr(), ev.data_ptr(), int(xyzt.shape[0]), dep.n_events, 1, 0, 100_000,
           C.byref(em), maps.data_ptr(), tmin.data_ptr(), npe.data_ptr(), None)
    assert rc == -1 and "device arrays" in lib.crysp_last_error().decode()
