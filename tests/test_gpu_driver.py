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
