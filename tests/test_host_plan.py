"""Host-side solver set-up without a device.

gfx_solver_create builds its whole host plan (element order, node numbering
with the C-PML direction-mask grouping, nodal C-PML coefficients, gather
slots, recording tables) and checks its invariants before the first device
call, so on a machine with no HIP device the call must end in
GFX_ERR_DEVICE ("no HIP device available"), never in an internal check or a
crash.  A numbering that is not a permutation of the nodes (a node queued
once per element that touches it) fails the plan's own check instead."""
import numpy as np
import pytest

import meshcases as mc
import pygfx

GFX_ERR_DEVICE = 2


@pytest.fixture(scope="module", autouse=True)
def _no_device():
    if pygfx.device_count() > 0:
        pytest.skip("a HIP device is visible: the device-free path is not reached")


@pytest.mark.parametrize("assembly", [pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_ATOMIC,
                                      pygfx.GFX_ASSEMBLY_GATHER])
@pytest.mark.parametrize("has_cpml", [True, False])
def test_create_plans_fully_before_the_device_check(assembly, has_cpml):
    mesh = mc.small_layered(nx=5, ny=4, nz=4, pml=1)
    with pytest.raises(pygfx.GfxError) as ei:
        pygfx.Solver(mesh, n_fields=3, assembly=assembly, has_cpml=has_cpml, record_capacity=2)
    assert ei.value.status == GFX_ERR_DEVICE, str(ei.value)
    assert "no HIP device" in str(ei.value)


def test_create_still_rejects_bad_arguments_first():
    mesh = mc.small_halfspace(nx=4, ny=4, nz=3, pml=1)
    with pytest.raises(pygfx.GfxError, match="elems_per_block"):
        pygfx.Solver(mesh, elems_per_block=3)
    with pytest.raises(pygfx.GfxError, match="pair element kernel"):
        pygfx.Solver(mesh, n_fields=2, element_kernel=pygfx.GFX_ELEMENT_PAIR, elems_per_block=8)
    with pytest.raises(pygfx.GfxError, match="elems_per_block must be 0"):
        pygfx.Solver(mesh, element_kernel=pygfx.GFX_ELEMENT_PENCIL, elems_per_block=2)
    with pytest.raises(pygfx.GfxError, match="unknown element kernel"):
        pygfx.Solver(mesh, element_kernel=5)
    good = mesh.arrays["ibool"]
    bad = np.array(good, copy=True)
    bad.reshape(-1)[7] = mesh.n_node
    mesh.arrays["ibool"] = bad
    try:
        with pytest.raises(pygfx.GfxError, match="ibool"):
            pygfx.Solver(mesh)
    finally:
        mesh.arrays["ibool"] = good


@pytest.mark.parametrize("assembly", [pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_GATHER])
def test_complete_cpml_plans_and_rejects_bad_inputs(assembly):
    mesh = mc.small_layered(nx=5, ny=4, nz=4, pml=1, want_profiles=1, pml_profile=1)
    with pytest.raises(pygfx.GfxError) as ei:
        pygfx.Solver(mesh, n_fields=2, assembly=assembly, cpml_complete=True, record_capacity=2)
    assert ei.value.status == GFX_ERR_DEVICE, str(ei.value)
    with pytest.raises(pygfx.GfxError, match="needs has_cpml"):
        pygfx.Solver(mesh, cpml_complete=True, has_cpml=False)
    bad = mesh.derive(pml_K=np.where(np.arange(mesh.pml_K.size) == 3 * 125 * 2 + 1, 0.0, mesh.pml_K))
    with pytest.raises(pygfx.GfxError, match="pml_K must be > 0"):
        pygfx.Solver(bad, cpml_complete=True)
    bad = mesh.derive(pml_alpha=-mesh.pml_alpha)
    with pytest.raises(pygfx.GfxError, match="pml_alpha >= 0"):
        pygfx.Solver(bad, cpml_complete=True)
    no_profiles = mc.small_layered(nx=5, ny=4, nz=4, pml=1)
    with pytest.raises(ValueError, match="want_profiles"):
        pygfx.Solver(no_profiles, cpml_complete=True)
