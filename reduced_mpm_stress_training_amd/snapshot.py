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


def frame_dict(x0, u, v, F_disp, stress, C, mass, time, friction_angle=None) -> dict:
    """The datasets of one frame (mpm_solver_warp.py:350-403) from (N, ...)
    arrays: x, q, v (3, N); time (1, 1); masses (1, N); f_tensor, C (9, N)
    row-major; stress (6, N) Voigt rows [xx, xy, xz, yy, yz, zz] (:383);
    mu (1, N) = the friction angle when given."""
    n = int(np.shape(x0)[0])
    out = {
        "x": np.asarray(x0).T,
        "q": np.asarray(u).T,
        "time": np.array([time]).reshape(1, 1),
        "masses": np.asarray(mass).reshape(1, n),
        "f_tensor": np.asarray(F_disp).reshape(-1, 9).T,
        "stress": np.asarray(stress).reshape(-1, 9)[:, VOIGT].T,
        "v": np.asarray(v).T,
        "C": np.asarray(C).reshape(-1, 9).T,
    }
    if friction_angle is not None:
        out["mu"] = np.ones((1, n)) * friction_angle
    return out


def frame_arrays(sim, include_parameter="friction_angle") -> dict:
    u, v, F_disp, stress = sim.export_uvFstress_from_warp()
    fa = sim.mpm_model.friction_angle if include_parameter == "friction_angle" else None
    return frame_dict(sim.export_x_initial().cpu().numpy(), u.cpu().numpy(), v.cpu().numpy(),
                      F_disp.cpu().numpy(), stress.cpu().numpy(), sim.export_C().cpu().numpy(),
                      sim.export_mass().cpu().numpy(), sim.time, fa)


def stress_from_voigt(st6: np.ndarray) -> np.ndarray:
    """(N, 6) Voigt rows [xx, xy, xz, yy, yz, zz] -> symmetric (N, 3, 3), as
    load_uvFstress_from_h5 rebuilds it (mpm_solver_warp.py:285-294)."""
    st6 = np.asarray(st6)
    out = np.zeros((st6.shape[0], 3, 3))
    for k, (i, j) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        out[:, i, j] = st6[:, k]
        out[:, j, i] = st6[:, k]
    return out


def f_tensor_rows(f, n: int) -> np.ndarray:
    """A sampling file's f_tensor as (n, 9) row-major: the reference reads it
    as (n, 3, 3) (mpm_solver_warp.py:136, 148); frames store (9, n)."""
    f = np.asarray(f)
    if f.shape == (9, n) and n != 9:
        return f.T.reshape(n, 9)
    return f.reshape(n, 9)


def resolve_frame_path(path: str) -> str:
    """The file to read for a frame or sampling path: the path itself, or its
    .npz twin when the .h5 name is given and h5py is not installed."""
    if os.path.exists(path) and (h5py is not None or not path.endswith(".h5")):
        return path
    if path.endswith(".h5") and os.path.exists(path[:-3] + ".npz"):
        return path[:-3] + ".npz"
    raise FileNotFoundError(f"h5 file cannot be found at {path}")


def write_sampling(path, x, masses, particle_volume, particle_density, f_tensor=None):
    """A sampling file in the layout load_from_sampling reads
    (mpm_solver_warp.py:134-143): x, q (3, N), masses, particle_volume,
    particle_density (1, N), f_tensor (N, 3, 3).  .npz (with these dataset
    names) unless the path ends in .h5 and h5py is installed."""
    x = np.asarray(x, dtype=np.float32)
    n = x.shape[0]
    arrs = {"x": x.T, "q": np.zeros((3, n), np.float32),
            "masses": np.asarray(masses, np.float32).reshape(1, n),
            "particle_volume": np.asarray(particle_volume, np.float32).reshape(1, n),
            "particle_density": np.asarray(particle_density, np.float32).reshape(1, n),
            "f_tensor": (np.zeros((n, 3, 3), np.float32) if f_tensor is None
                         else np.asarray(f_tensor, np.float32).reshape(n, 3, 3))}
    if path.endswith(".h5") and h5py is not None:
        with h5py.File(path, "w") as f:
            for k, a in arrs.items():
                f.create_dataset(k, data=a)
    else:
        np.savez(path[:-3] + ".npz" if path.endswith(".h5") else path, **arrs)


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
        # the reference's messages (mpm_solver_warp.py:382, 404)
        print("stress out:", np.linalg.norm(sim.export_uvFstress_from_warp()[3].reshape(-1, 9).cpu().numpy()))
        if h5py is not None:
            path = stem + ".h5"
            if os.path.exists(path):
                os.remove(path)
            with h5py.File(path, "w") as f:
                for k, a in arrs.items():
                    f.create_dataset(k, data=a)
        else:
            path = stem + ".npz"
            np.savez(path, **arrs)
        print("save siumlation data at frame ", frame, " to ", path)


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
    return read_frame(resolve_frame_path(path))
