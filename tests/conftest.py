import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def workload_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("workloads"))


@pytest.fixture(scope="session")
def cornell(workload_dir):
    from luz_amd import workloads
    return workloads.write_cornell(os.path.join(workload_dir, "cornell"))


@pytest.fixture(scope="session")
def spheres(workload_dir):
    from luz_amd import workloads
    return workloads.spheres_scene(os.path.join(workload_dir, "spheres"))


@pytest.fixture(scope="session")
def primitives(workload_dir):
    from luz_amd import workloads
    return workloads.write_primitives(os.path.join(workload_dir, "primitives"))


@pytest.fixture(scope="session")
def materials(workload_dir):
    from luz_amd import workloads
    return workloads.write_materials(os.path.join(workload_dir, "materials"))


@pytest.fixture(scope="session")
def meshlight(workload_dir):
    from luz_amd import workloads
    return workloads.write_meshlight(os.path.join(workload_dir, "meshlight"))


@pytest.fixture(scope="session")
def skylinear(workload_dir):
    from luz_amd import workloads
    return workloads.write_skylinear(os.path.join(workload_dir, "skylinear"))


@pytest.fixture(scope="session")
def atmosphere(workload_dir):
    from luz_amd import workloads
    return workloads.write_atmosphere(os.path.join(workload_dir, "atmosphere"))


@pytest.fixture(scope="session")
def volume(workload_dir):
    from luz_amd import workloads
    return workloads.write_volume(os.path.join(workload_dir, "volume"))


@pytest.fixture(scope="session")
def ctx():
    from luz_amd import Context
    c = Context(0)
    yield c
    c.close()


def rel_mse(x, r):
    """Relative MSE of an image against a reference: sum((x-r)^2) / sum(r^2)."""
    import numpy as np
    x = x.astype(np.float64)
    r = r.astype(np.float64)
    return float(((x - r) ** 2).sum() / max((r ** 2).sum(), 1e-300))


@pytest.fixture(scope="session")
def spectral(workload_dir):
    from luz_amd import workloads
    return workloads.write_spectral(os.path.join(workload_dir, "spectral"))


@pytest.fixture(scope="session")
def atmosphere_env(workload_dir):
    from luz_amd import workloads
    return workloads.write_atmosphere_env(os.path.join(workload_dir, "atmosphere_env"))


@pytest.fixture(scope="session")
def optics(workload_dir):
    from luz_amd import workloads
    return workloads.write_optics(os.path.join(workload_dir, "optics"))


@pytest.fixture(scope="session")
def subsurface(workload_dir):
    from luz_amd import workloads
    return workloads.write_subsurface(os.path.join(workload_dir, "subsurface"))


@pytest.fixture(scope="session")
def invisible(workload_dir):
    from luz_amd import workloads
    return workloads.write_invisible(os.path.join(workload_dir, "invisible"))


@pytest.fixture(scope="session")
def ies(workload_dir):
    from luz_amd import workloads
    return workloads.write_ies(os.path.join(workload_dir, "ies"))


@pytest.fixture(scope="session")
def sunlit(workload_dir):
    from luz_amd import workloads
    return workloads.write_sunlit(os.path.join(workload_dir, "sunlit"))


@pytest.fixture(scope="session")
def terrain(workload_dir):
    """The volume workload's scene: 1,002,528-triangle terrain, haze, sun (rendered at
    reduced sizes by the tests)."""
    from luz_amd import workloads
    return workloads.terrain_scene(os.path.join(workload_dir, "terrain"))


@pytest.fixture(scope="session")
def batch(workload_dir):
    """The uneven 32-scene batch (rendered at reduced sizes by the tests)."""
    from luz_amd import workloads
    return workloads.batch_scenes(os.path.join(workload_dir, "batch"))
