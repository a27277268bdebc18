"""The oracle's FK injection (incident field through the Stacey faces) against the
closed-form vertically incident P plane wave, and against a zero incident field."""
import numpy as np

from oracle import oracle as orc
from fk_plane_wave import ALPHA, BETA, RHO, incident_fields, plane_wave


def _oracle(nstep, dt, depths, nex=4, ney=4, nez=16, h=500.0):
    n = nex * ney * nez * 125
    return orc.Oracle(nex, ney, nez, nex * h, ney * h, nez * h, np.full(n, RHO, np.float32),
                      np.full(n, ALPHA, np.float32), np.full(n, BETA, np.float32), dt=dt, nstep=nstep,
                      src_xyz=[[1000.0, 1000.0, -4000.0]], src_mt=[[0.0] * 6], src_f0=[1.0],
                      rec_xyz=[[1000.0, 1000.0, z] for z in depths], precision=64)


def _face_fields(o, nstep, dt, L):
    f = o.faces()
    mesh = o.mesh()
    ib = mesh["ibool"].reshape(-1, 125)
    loc = (f["ijk"][..., 0] + 5 * f["ijk"][..., 1] + 25 * f["ijk"][..., 2])
    z = mesh["xyz"][ib[f["ispec"][:, None], loc], 2].ravel()
    return incident_fields(z, f["normal"], nstep, dt, L)


def test_oracle_injection_reproduces_the_plane_wave():
    depths = [0.0, -1500.0, -3000.0, -5000.0, -6500.0]
    nstep, dt, L = 1125, 0.004, 8000.0
    o = _oracle(nstep, dt, depths)
    assert o.faces()["ispec"].min() >= 0  # 0-based element indices
    V, T = _face_fields(o, nstep, dt, L)
    o.set_fk(V, T)
    seis = o.forward(save=False)
    t = np.arange(nstep) * dt
    want = np.stack([plane_wave(z, t, L)[0] for z in depths], axis=1)
    err = np.linalg.norm(seis[:, :, 2] - want) / np.linalg.norm(want)
    late = np.linalg.norm(seis[1:, :, 2] - want[:-1]) / np.linalg.norm(want)
    assert err < 1e-3 and late > 5 * err, (err, late)
    assert np.abs(seis[:, :, :2]).max() < 1e-5 * np.abs(want).max()
    # linear interpolation of a series stored every 4th step stays close
    o.set_fk(V[::4], T[::4], resamp=4)
    err4 = np.linalg.norm(o.forward(save=False)[:, :, 2] - want) / np.linalg.norm(want)
    assert err4 < 3e-3
    # a zero incident field leaves the box at rest
    o.set_fk(0 * V, 0 * T)
    assert np.abs(o.forward(save=False)).max() == 0.0
