"""CPU-side checks of the C-ABI library and the host logic (no kernel runs)."""
import ctypes as Ct
import os
import re

import pytest

from reduced_mpm_stress_training_amd import _lib as L
from reduced_mpm_stress_training_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    names = set()
    for hdr, prefix in (("mpm_hip.h", "mpm_"), ("crom_hip.h", "crom_")):
        src = open(os.path.join(ROOT, "include", hdr)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = L.load()
    declared = _header_functions()
    assert declared, "no functions parsed from include/*.h"
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(L.exported_symbols()) == declared


def test_abi_version_and_errors():
    lib = L.load()
    assert lib.mpm_abi_version() == L.MPM_ABI_VERSION
    prm = L.MpmParams()
    prm.n_grid = 2
    prm.material = 2
    st = L.MpmState()
    assert lib.mpm_p2g(prm, st, None, 1e-4, None) < 0
    assert b"n_grid" in lib.mpm_last_error()
    prm.n_grid = 32
    prm.material = 3
    assert lib.mpm_g2p(prm, st, None, 1e-4, None) < 0
    assert b"material" in lib.mpm_last_error()
    prm.material = 1
    prm.hardening = 2
    assert lib.mpm_g2p(prm, st, None, 1e-4, None) < 0
    assert b"hardening" in lib.mpm_last_error()
    prm.hardening = 0
    # the ordered-state (adjoint) path is sand only
    assert lib.mpm_step_ordered(prm, st, None, None, 0, None, 0, 1e-4, 0.0, None) < 0
    assert b"sand" in lib.mpm_last_error()
    prm.material = 2
    assert lib.mpm_step(prm, st, None, 0, None, 0, 0.0, 1e-4, None) < 0
    st.n, st.stride, st.planes = 10, 10, 1
    assert lib.mpm_step(prm, st, None, 0, None, 0, 0.0, 1e-4, None) < 0
    assert b"workspace" in lib.mpm_last_error()
    cols = (L.MpmCollider * 1)()
    assert lib.mpm_grid_update(prm, 1, 1, cols, 17, 0.0, 1e-4, None) < 0
    assert lib.mpm_workspace_bytes(prm, 1000) >= 2 * 32 ** 3 * 16


def test_material_parameters_follow_reference():
    mu, lam = S.lame_from_E_nu(2e5, 0.3)
    assert abs(mu - 2e5 / 2.6) < 1e-9 and abs(lam - 2e5 * 0.3 / 1.3 / 0.4) < 1e-6
    a = S.alpha_from_friction_angle(30.0)
    assert abs(a - (2.0 / 3.0) ** 0.5 * 2 * 0.5 / 2.5) < 1e-8


def test_solver_host_logic_on_cpu_tensors():
    from reduced_mpm_stress_training_amd.mpm_solver import MPM_Simulator_HIP, MPM_Simulator_WARP
    assert MPM_Simulator_WARP is MPM_Simulator_HIP
    sim = MPM_Simulator_HIP(10, n_grid=16, device="cpu")
    with pytest.raises(TypeError):
        sim.set_parameters(material="water")
    with pytest.raises(ValueError):
        sim.add_surface_collider((0, 0, 0), (0, 2, 0), "sticky", friction=0.1)
    sim.set_parameters(material="sand", friction_angle=30, g=9.8, mu=1e5, lam=0.3)
    assert sim.mpm_model.material == 2 and sim._gravity_vec() == (0.0, -9.8, 0.0)
    sim.add_surface_collider((0, 0.58, 0), (0, 2, 0), "sticky", 0.0)
    assert sim.colliders[0]["normal"] == (0.0, 1.0, 0.0)
    sim.add_bounding_box()
    with pytest.raises(FileNotFoundError):
        sim.load_from_sampling("/does/not/exist.h5")


def test_segment_entry_points_reject_bad_arguments():
    """mpm_record_segment / mpm_adjoint_recorded validate their arguments on
    the host before anything is launched (the pointers here are never
    dereferenced)."""
    lib = L.load()
    prm = L.MpmParams()
    prm.n_grid, prm.material, prm.dx, prm.inv_dx = 32, 2, 1.0 / 32, 32.0
    cols = (L.MpmCollider * 1)()
    st = (L.MpmState * 2)()
    for k in range(2):
        st[k].n, st[k].stride, st[k].planes = 100, 128, 0x100000 * (k + 1)
    grids = (Ct.c_void_p * 2)(0x400000, 0x800000)
    order, grid_bytes = 0x1000000, lib.mpm_step_grids_bytes(prm)
    # accumulator and velocities (32^3 float4 each), 4^3 block marks padded to 256 bytes
    assert grid_bytes == 2 * 16 * 32 ** 3 + 256
    assert lib.mpm_adjoint_workspace_bytes(prm) == grid_bytes + 2 * 16 * 32 ** 3
    assert lib.mpm_run_ordered_workspace_bytes(prm) == 2 * grid_bytes
    last = L.MpmState()
    last.n, last.stride, last.planes = 100, 128, st[1].planes
    rec = lambda n_sub, last_p, nxt, nbytes: lib.mpm_record_segment(prm, st, order, grids, cols, 0, 0, 1e-4, n_sub, 0,
                                                                    last_p, nxt, nbytes, None)
    assert rec(0, None, None, 0) < 0                                      # n_sub < 1
    assert rec(2, Ct.byref(last), 0x2000000, grid_bytes) < 0              # last aliases a state
    assert b"aliases a state" in lib.mpm_last_error()
    last.planes = 0x200000 * 3
    assert rec(2, Ct.byref(last), grids[1], grid_bytes) < 0               # next grid is a step grid
    assert b"aliases a step grid" in lib.mpm_last_error()
    assert rec(2, Ct.byref(last), 0x2000000, grid_bytes - 16) < 0         # next grid too small
    assert rec(2, Ct.byref(last), 0x2000010, grid_bytes) < 0              # next grid not 256-byte aligned
    last.stride = 256
    assert rec(2, Ct.byref(last), 0x2000000, grid_bytes) < 0              # last differs in stride
    st[1].stride = 256
    assert rec(2, None, None, 0) < 0                                      # states differ in stride
    st[1].stride = 128
    grids[1] = 0x800010
    assert rec(2, None, None, 0) < 0                                      # step grid not aligned
    grids[1] = 0x800000
    adj = lib.mpm_adjoint_recorded
    assert adj(prm, st, order, grids, cols, 0, 0x3000000, 0, 1e-4, 2, 0x4000000, 0x4000000, 128, 0x5000000, None) < 0
    assert b"adjoint buffers" in lib.mpm_last_error()                     # adj_a == adj_b


def test_bins_reject_oversized_bins_and_size_the_tiles():
    """The binned entry points check MpmBins on the host before any launch
    (the pointers here are never dereferenced): more than 16e6 slots per bin
    is refused (the fused kernel's split arithmetic counts passes of 512 in
    15 bits).  The workspace holds, per bin, a float4 tile of 10^3 nodes and
    the block lists (the parts of split bins add into the FP64 spill grid),
    after the spill grid and the velocity grid."""
    lib = L.load()
    prm = L.MpmParams()
    prm.n_grid, prm.material, prm.dx, prm.inv_dx = 32, 2, 1.0 / 32, 32.0
    prm.mu, prm.lam, prm.alpha = 7.7e4, 1.15e5, 0.2
    nb = 4
    bins = L.MpmBins()
    bins.nb, bins.bin_cap, bins.overflow_cap = nb, 16_100_000, 16_100_000
    bins.stride = (nb ** 3 * bins.bin_cap + bins.overflow_cap) * L.NPLANES
    bins.planes[0], bins.planes[1], bins.count = 0x1000000, 0x2000000, 0x3000000
    assert lib.mpm_bins_g2p2g(prm, bins, 0x4000000, 1e-4, None) < 0
    assert b"16e6 slots per bin" in lib.mpm_last_error()
    # the spill grid (4 doubles per node) first, then the velocity grid (float4)
    nodes = 32 ** 3
    assert lib.mpm_bins_velocity_offset(prm) == nodes * 32
    assert lib.mpm_bins_status_offset(prm) >= nodes * 32 + nodes * 16
    assert lib.mpm_bins_workspace_bytes(prm) >= nb ** 3 * (10 ** 3 * 16 + 4 * 4) + nodes * (32 + 16)
