"""Pin the oracle's prescribed bottom layer (if_prescribed_m,
magnetic_solver.cpp:1526-1609 and 1891-1957; start-up
magnetic_initialize.cpp:1074-1086):

- at start-up and after every stage the ferromagnetic cells of k = 0 sit
  on the cone (sqrt(1 - m0^2), m0 sin wt, m0 cos wt), m0 = sin(angle),
  w = 2 pi f, at the times the reference's updates use: the stage-2 and
  stage-3 stores at t - dt/2 (update_m_RK1 and _RK2 call the half-step
  form), the stage-4 store at t (update_m_RK3, the full-step form) and m
  after the step at t (update_m); the rest of the film, vacuum and
  non-magnetic cells of the bottom layer follow the plain RK4 update;
- driven this way the layer pumps a spin wave at frequency f into the film
  above it: the transverse m of the next layer oscillates at f.

A planted mistake (GFO_PLANT = 8: the stage-2 update writes the cone at t
instead of t - dt/2) fails the stage-time check.
"""
import numpy as np
import pytest

from goferro import FIELD_M, FIELD_M_STORE, FLAG_PRESCRIBED_M, Material, Problem
from goferro.workloads import GYRO, tilted_m

ANGLE, FREQ = 12.0, 5e10


def cone(t, angle=ANGLE, freq=FREQ):
    """The prescribed direction at time t, from the angle in degrees and
    the frequency in Hz (the textbook cone about x)."""
    th = np.deg2rad(angle)
    w = 2 * np.pi * freq
    return np.array([np.cos(th), np.sin(th) * np.sin(w * t), np.sin(th) * np.cos(w * t)])


def film(shape=(4, 3, 5), dt=2e-14, hole=True):
    """A ferromagnetic film with a vacuum cell and a non-magnetic cell in
    its bottom layer."""
    fm = Material.make(if_FM=1, anisotropy_type=2, Ms=8e5, gyro=GYRO, FM_damping=0.05, K1=2e4, Aex=1e-11)
    dielectric = Material.make(r_permittivity=4.0)
    cell = np.ones(shape, np.uint32)
    if hole:
        cell[0, 0, 0] = 0
        cell[1, 2, 0] = 2
    prob = Problem(shape=shape, spacing=(2e-9,) * 3, materials=[fm, dielectric], material_cell=cell, dt=dt,
                   periodic=(True, True, False), flags=FLAG_PRESCRIBED_M, Hext_stat=(0.0, 0.0, 2e4),
                   precess_angle=ANGLE, precess_frequency=FREQ)
    return prob, cell


def staged_bottoms(oracle_lib, lib_path=None, nsteps=3):
    """Run nsteps steps stage by stage; return the bottom-layer store after
    every stage and m after every step, with their times."""
    prob, cell = film()
    o = oracle_lib.Oracle(prob, lib_path)
    m0 = tilted_m(prob.shape, 7, 0.3)
    o.set(FIELD_M, m0)
    o.prescribe_m()
    out = [("start", 0.0, o.get(FIELD_M)[:, :, :, 0], o.get(FIELD_M_STORE)[:, :, :, 0])]
    for _ in range(nsteps):
        o.time_advance()
        t = o.time
        for s in (1, 2, 3, 4):
            o.mag_stage(s)
            ts = t - prob.dt * 0.5 if s <= 2 else t
            out.append((f"stage {s}", ts, o.get(FIELD_M_STORE)[:, :, :, 0], None))
        out.append(("step", t, o.get(FIELD_M)[:, :, :, 0], None))
    o.close()
    return cell, m0, out


def check_cones(cell, m0, rows):
    fm = cell[:, :, 0] == 1
    for what, t, plane, store in rows:
        want = cone(t)
        for a in range(3):
            np.testing.assert_allclose(plane[a][fm], want[a], rtol=0, atol=1e-15, err_msg=what)
        if store is not None:
            np.testing.assert_array_equal(store, plane)
    # the vacuum and dielectric cells of the bottom layer are never written
    last = rows[-1][2]
    for i, j in ((0, 0), (1, 2)):
        np.testing.assert_array_equal(last[:, i, j], m0[:, i, j, 0])


def test_prescribed_layer_at_stage_times(oracle_lib):
    cell, m0, rows = staged_bottoms(oracle_lib)
    check_cones(cell, m0, rows)


def test_prescribed_layer_pin_fails_on_planted_time(oracle_lib, tmp_path):
    path = oracle_lib.build_planted(8, tmp_path)
    cell, m0, rows = staged_bottoms(oracle_lib, path)
    with pytest.raises(AssertionError):
        check_cones(cell, m0, rows)


def test_rest_of_film_follows_plain_rk4(oracle_lib):
    """Cells above the bottom layer see the prescribed layer only through
    the exchange stencil: with the layer's cone taken as the initial state
    of an unprescribed run at t = 0 the first stage-1 increments agree
    everywhere above k = 1 (the stencil reaches one layer)."""
    prob, cell = film(hole=False)
    m0 = tilted_m(prob.shape, 3, 0.2)
    a = oracle_lib.Oracle(prob)
    a.set(FIELD_M, m0)
    a.prescribe_m()
    start = a.get(FIELD_M)
    bare = Problem(**{**prob.__dict__, "flags": 0})
    b = oracle_lib.Oracle(bare)
    b.set(FIELD_M, start)
    a.time_advance()
    b.time_advance()
    a.mag_stage(1)
    b.mag_stage(1)
    sa, sb = a.get(FIELD_M_STORE), b.get(FIELD_M_STORE)
    np.testing.assert_array_equal(sa[:, :, :, 1:], sb[:, :, :, 1:])
    assert np.abs(sa[:, :, :, 0] - sb[:, :, :, 0]).max() > 0
    a.close()
    b.close()


def test_layer_pumps_its_frequency_into_the_film(oracle_lib):
    """The exchange-coupled layer above the prescribed one oscillates at
    the prescribed frequency (peak of its m_y spectrum within one bin)."""
    shape, dt, f = (1, 1, 6), 2e-14, 2e10
    fm = Material.make(if_FM=1, anisotropy_type=0, Ms=8e5, gyro=GYRO, FM_damping=0.2, Aex=1e-11)
    prob = Problem(shape=shape, spacing=(2e-9,) * 3, materials=[fm], material_cell=np.ones(shape, np.uint32), dt=dt,
                   periodic=(True, True, False), flags=FLAG_PRESCRIBED_M, Hext_stat=(2e5, 0.0, 0.0),
                   precess_angle=20.0, precess_frequency=f)
    o = oracle_lib.Oracle(prob)
    m = np.zeros((3,) + shape)
    m[0] = 1.0
    o.set(FIELD_M, m)
    o.prescribe_m()
    n_per = int(round(1 / (f * dt)))
    series = []
    for _ in range(8 * n_per):
        o.step(1)
        series.append(o.get(FIELD_M)[1, 0, 0, 1])
    y = np.asarray(series[2 * n_per:])
    spec = np.abs(np.fft.rfft(y - y.mean()))
    freqs = np.fft.rfftfreq(y.size, dt)
    assert abs(freqs[np.argmax(spec)] - f) <= freqs[1], (freqs[np.argmax(spec)], f)
    o.close()
