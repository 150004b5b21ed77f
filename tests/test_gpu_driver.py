"""The drop-in driver: `apcloud file.param -ksp_*` (reference argv
convention, test.cpp:16-80) reading the reference's .param and ASCII VTK IC
formats and running the C++ Controller/Mover/Solver/Octree over the device,
checked against the oracle's restatement of Controller::solve."""
import os
import subprocess

import numpy as np
import pytest

import apcloud
import pyoracle as po
from apcloud import ic

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(apcloud.LIB_PATH), "apcloud")


def _params(tmp, n, final, steps_end, extra=None):
    tags = dict(
        Omega=1.0, OmegaLambda=0.0, Redshift=49.0, HubbleParam=0.7, GravitionalConstant=1.0,
        UnitLength_in_cm=1.0, UnitMass_in_g=1.0, UnitVelocity_in_cm_per_s=1.0,
        ICFile=os.path.join(tmp, "ic.vtk"), ICFileType=1, FileBase="apc",
        OutputFile=str(tmp), BoundingBox_min_x=0.0, BoundingBox_max_x=1.0,
        BoundingBox_min_y=0.0, BoundingBox_max_y=1.0, BoundingBox_min_z=0.0,
        BoundingBox_max_z=1.0, Periodic=1, DeltaA=0.004, EndRedShift=steps_end,
        MassPerParticle=1.0 / n, MaxOctreeLevel=20, CellSizeParameter=0.0, NumberOfNeighbors=27,
        DensityEstimator=1, Units=2, AdaptiveTimeStep=0, LeafCapacity=8, WriteInterval=0,
        FinalStateFile=final)
    if extra:
        tags.update(extra)
    path = os.path.join(tmp, "run.param")
    ic.write_param(path, **tags)
    return path


def test_driver_runs_controller_like_the_oracle(tmp_path):
    assert os.path.exists(EXE), "driver not built"
    x, y, z, vx, vy, vz = ic.zeldovich(16, seed=3, rms=0.01)
    n = len(x)
    ic.write_vtk(os.path.join(tmp_path, "ic.vtk"), x, y, z, vx, vy, vz)
    final = os.path.join(tmp_path, "final.bin")
    end_z = 1.0 / 0.032 - 1.0  # a = 0.02 .. 0.032: 3 steps of 0.004
    param = _params(str(tmp_path), n, final, end_z)
    print("case: driver n=%d" % n, flush=True)
    p = subprocess.run([EXE, param, "-ksp_rtol", "1e-12", "-ksp_gmres_restart", "60",
                        "-ksp_max_it", "3000"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    got = np.fromfile(final, dtype=np.float64).reshape(9, n)
    # the VTK round trip is exact (17 digits)
    o = po.Oracle(po.make_params(depth=20, bc=1, c=0.0, lm=27, leaf_capacity=8, units=2,
                                 omega=1.0, mass_per_particle=1.0 / n, rtol=1e-12, restart=60,
                                 max_it=3000))
    o.set_particles(x, y, z, vx, vy, vz)
    _, steps = o.run(1.0 / 50.0, 1.0 / (1.0 + end_z), 0.004)
    assert steps == 3
    assert p.stdout.count("Number of computational particles") == 3
    want = np.stack(o.particles())
    assert np.abs(got[:3] - want[:3]).max() < 1e-11
    rel = np.linalg.norm(got[6:] - want[6:]) / np.linalg.norm(want[6:])
    assert rel < 1e-6, rel


def test_driver_rejects_missing_tag(tmp_path):
    path = os.path.join(tmp_path, "bad.param")
    ic.write_param(path, Omega=1.0)
    p = subprocess.run([EXE, path], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
    assert "I miss a value for tag 'OmegaLambda'" in p.stdout


def test_driver_adaptive_time_step_on_a_lattice(tmp_path):
    """The reference's default build has ADAPTIVE_TIME_STEP on
    (apcloud_macro.h, controller.cpp:146-162).  On an 8^3 lattice at rest the
    step is dt = 0.05 Delta_A / sqrt(512) every step (test_oracle_pins.py
    derives it): 19 steps from a = 0.02 to 0.0254 with Delta_A = 0.01, the
    lattice unmoved, as the oracle's controller."""
    from test_oracle_pins import adaptive_lattice_case
    x, y, z, n, da, start, end, dt, steps = adaptive_lattice_case()
    zero = np.zeros(n)
    ic.write_vtk(os.path.join(tmp_path, "ic.vtk"), x, y, z, zero, zero, zero)
    final = os.path.join(tmp_path, "final.bin")
    end_z = 1.0 / end - 1.0
    param = _params(str(tmp_path), n, final, end_z,
                    extra=dict(AdaptiveTimeStep=1, DeltaA=da, Redshift=1.0 / start - 1.0))
    print("case: driver adaptive n=%d" % n, flush=True)
    p = subprocess.run([EXE, param, "-ksp_rtol", "1e-10", "-ksp_gmres_restart", "60",
                        "-ksp_max_it", "500"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert p.stdout.count("Number of computational particles") == steps
    got = np.fromfile(final, dtype=np.float64).reshape(9, n)
    assert max(np.abs(got[k] - p0).max() for k, p0 in enumerate((x, y, z))) < 1e-12
    o = po.Oracle(po.make_params(depth=20, bc=1, c=0.0, lm=27, leaf_capacity=8, units=2,
                                 omega=1.0, mass_per_particle=1.0 / n, rtol=1e-10, restart=60,
                                 max_it=500, adaptive_dt=1))
    o.set_particles(x, y, z, zero, zero, zero)
    _, done = o.run(1.0 / (1.0 + (1.0 / start - 1.0)), 1.0 / (1.0 + end_z), da)
    assert done == steps


def test_driver_writes_the_reference_snapshots(tmp_path):
    """WriteInterval 2 over three steps: the limit snapshot at a = 0.02, then
    both snapshots at steps 0 and 2 and at the end (controller.cpp:95-96,
    121-130, 222-223); the final physical snapshot byte for byte equals the
    reference's viewer.cpp on the final particle state."""
    x, y, z, vx, vy, vz = ic.zeldovich(16, seed=3, rms=0.01)
    n = len(x)
    ic.write_vtk(os.path.join(tmp_path, "ic.vtk"), x, y, z, vx, vy, vz)
    out = os.path.join(tmp_path, "out")
    os.makedirs(out)
    final = os.path.join(tmp_path, "final.bin")
    end_z = 1.0 / 0.032 - 1.0
    param = _params(str(tmp_path), n, final, end_z, extra=dict(WriteInterval=2, OutputFile=out))
    print("case: driver snapshots n=%d" % n, flush=True)
    p = subprocess.run([EXE, param, "-ksp_rtol", "1e-10", "-ksp_gmres_restart", "60",
                        "-ksp_max_it", "3000"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    want = ["physical_%f.vtk" % t for t in (0.02, 0.022, 0.030, 0.032)] + \
        ["computational_%f.vtk" % t for t in (0.022, 0.030, 0.032)]
    assert sorted(os.listdir(out)) == sorted(want)
    for nm in want[1:4]:
        lines = open(os.path.join(out, nm)).read().split("\n")
        # positions, velocity vectors and three scalars, acceleration vectors
        # and three scalars: 9 n data lines plus the block headers
        assert lines[4] == "POINTS %d double" % n and 9 * n < len(lines) < 9 * n + 40
    if not po.ref_available():
        pytest.skip("oracle/_ref not built: the byte comparison needs the reference viewer")
    got = np.fromfile(final, dtype=np.float64).reshape(9, n)
    ref_dir = os.path.join(tmp_path, "ref")
    os.makedirs(ref_dir)
    one = dict(x=[0.0], y=[0.0], z=[0.0], np=[0], type=[0], first=[0], density=[0.0],
               potential=[0.0], p_x=[0.0], p_y=[0.0], p_z=[0.0])
    assert po.ref_view(ref_dir, 0.5, 0.032, (0, 1, 0, 1, 0, 1), list(got), one) == 0
    a = open(os.path.join(out, "physical_0.032000.vtk"), "rb").read()
    b = open(os.path.join(ref_dir, "physical_0.032000.vtk"), "rb").read()
    assert a == b


def test_driver_rejects_an_unsupported_preconditioner(tmp_path):
    x, y, z, vx, vy, vz = ic.zeldovich(8, seed=3, rms=0.01)
    n = len(x)
    ic.write_vtk(os.path.join(tmp_path, "ic.vtk"), x, y, z, vx, vy, vz)
    param = _params(str(tmp_path), n, os.path.join(tmp_path, "final.bin"), 1.0 / 0.024 - 1.0)
    p = subprocess.run([EXE, param, "-pc_type", "ilu"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode != 0
    assert "-pc_type ilu not supported" in p.stderr
