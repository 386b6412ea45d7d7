"""Snapshot I/O in the reference's frame format.

save_frame restates MPM_Simulator_WARP.save_data_at_frame
(engine_stress/mpm_solver_warp.py:339-404) and particle_position2obj
(:312-322): datasets x (3,N) initial position, q (3,N) displacement,
time (1,1), masses (1,N), f_tensor (9,N) row-major F_disp, stress (6,N)
Voigt [xx,xy,xz,yy,yz,zz] (columns [0,1,2,4,5,8] of the row-major 3x3, :383),
v (3,N), mu (1,N) friction angle, C (9,N).  Files: h5_f_<frame:010d>.h5 when
h5py is importable, else h5_f_<frame:010d>.npz with the same dataset names;
OBJ lines "v x y z c c c" with c = 1 - 1/color (:319).
"""
from __future__ import annotations

import os

import numpy as np

try:  # h5py is optional; the reference pins h5py==3.4.0 (requirements.txt)
    import h5py  # type: ignore
except Exception:  # pragma: no cover - depends on the image
    h5py = None

VOIGT = [0, 1, 2, 4, 5, 8]


def frame_arrays(sim, include_parameter="friction_angle") -> dict:
    n = sim.n_particles
    u, v, F_disp, stress = sim.export_uvFstress_from_warp()
    out = {
        "x": sim.export_x_initial().cpu().numpy().T,
        "q": u.cpu().numpy().T,
        "time": np.array([sim.time]).reshape(1, 1),
        "masses": sim.export_mass().cpu().numpy().reshape(1, n),
        "f_tensor": F_disp.reshape(-1, 9).cpu().numpy().T,
        "stress": stress.reshape(-1, 9).cpu().numpy()[:, VOIGT].T,
        "v": v.cpu().numpy().T,
        "C": sim.export_C().reshape(-1, 9).cpu().numpy().T,
    }
    if include_parameter == "friction_angle":
        out["mu"] = np.ones((1, n)) * sim.mpm_model.friction_angle
    return out


def write_obj(path, xyz: np.ndarray, color: np.ndarray):
    c = 1.0 - 1.0 / color
    lines = ["v %s %s %s %s %s %s" % (repr(float(p[0])), repr(float(p[1])), repr(float(p[2])),
                                      repr(float(ci)), repr(float(ci)), repr(float(ci)))
             for p, ci in zip(xyz, c)]
    with open(path, "w") as f:
        f.write("\n".join(lines))
        if lines:
            f.write("\n")


def save_frame(sim, dir_name, frame, include_parameter="friction_angle", save_to_h5=True, save_to_obj=True,
               save_visual_F=False):
    os.makedirs(dir_name, 0o777, exist_ok=True)
    stem = os.path.join(dir_name, "h5_f_" + str(frame).zfill(10))
    if save_to_obj:
        write_obj(stem + ".obj", sim.export_x().cpu().numpy(), sim.particle_color)
    if save_visual_F:
        F = sim.export_uvFstress_from_warp()[2].reshape(-1, 9).cpu().numpy()
        col = np.minimum(1.0, np.linalg.norm(F, axis=1) / 2.0)
        x0 = sim.export_x_initial().cpu().numpy()
        with open(os.path.join(dir_name, "visualF_f_" + str(frame).zfill(10) + ".obj"), "w") as f:
            for p, ci in zip(x0, col):
                f.write("v %r %r %r %r 0 0\n" % (float(p[0]), float(p[1]), float(p[2]), float(ci)))
    if save_to_h5:
        arrs = frame_arrays(sim, include_parameter)
        if h5py is not None:
            path = stem + ".h5"
            if os.path.exists(path):
                os.remove(path)
            with h5py.File(path, "w") as f:
                for k, a in arrs.items():
                    f.create_dataset(k, data=a)
        else:
            np.savez(stem + ".npz", **arrs)


def read_frame(path) -> dict:
    if path.endswith(".h5"):
        if h5py is None:
            raise ImportError("h5py is not installed; read the .npz twin instead")
        with h5py.File(path, "r") as f:
            return {k: f[k][()] for k in f.keys()}
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def read_sampling(path) -> dict:
    """Initial sampling file (mpm_solver_warp.py:134-143): datasets x, q (3,N),
    masses, particle_volume, particle_density (1,N), f_tensor."""
    return read_frame(path)
