"""Independent numpy Yee time stepping on global field arrays (test helper).

Written from the split-field Yee equations on the global component shapes (Ex (Nx,Ny+1,Nz+1) ...),
with no x slabs, halos or index shifts.  It consumes the CPU-packed coefficient dictionary of
rfnetwork/solver.py:1463-1523 (gpu=False) and the same probe/monitor dictionaries, and follows the
same per-element operation order as rfnetwork/core/src/solver.cpp (E: Ca*s + Cb*(h+ - h);
H: Da*s + (Db2*e - Db1*e-)), so in float32 its results must equal the slab-threaded oracle's bit for
bit.  Probes use global indices directly; monitors read global planes.
"""
import numpy as np


def run(coefficients, probes, monitors, Nx, Ny, Nz, Nt, dtype=np.float32):
    c = {k: v.astype(dtype, copy=True) for k, v in coefficients.items()}
    # PEC wall at x = Nx (the solver zeroes these coefficient planes)
    for k in ("Cb_ey_z", "Cb_ey_x", "Ca_ey_z", "Ca_ey_x", "Cb_ez_x", "Cb_ez_y", "Ca_ez_x", "Ca_ez_y",
              "Db_hx_y1", "Db_hx_y2", "Db_hx_z1", "Db_hx_z2", "Da_hx_y", "Da_hx_z"):
        c[k][Nx - 1] = 0
    z = lambda *s: np.zeros(s, dtype)
    f = dict(ex_y=z(Nx, Ny + 1, Nz + 1), ex_z=z(Nx, Ny + 1, Nz + 1), ey_z=z(Nx + 1, Ny, Nz + 1),
             ey_x=z(Nx + 1, Ny, Nz + 1), ez_x=z(Nx + 1, Ny + 1, Nz), ez_y=z(Nx + 1, Ny + 1, Nz),
             hx_y=z(Nx + 1, Ny, Nz), hx_z=z(Nx + 1, Ny, Nz), hy_z=z(Nx, Ny + 1, Nz), hy_x=z(Nx, Ny + 1, Nz),
             hz_x=z(Nx, Ny, Nz + 1), hz_y=z(Nx, Ny, Nz + 1))
    f["ex"], f["ey"], f["ez"] = z(Nx, Ny + 1, Nz + 1), z(Nx + 1, Ny, Nz + 1), z(Nx + 1, Ny + 1, Nz)
    f["hx"], f["hy"], f["hz"] = z(Nx + 1, Ny, Nz), z(Nx, Ny + 1, Nz), z(Nx, Ny, Nz + 1)
    names = ("ex", "ey", "ez", "hx", "hy", "hz")
    split = dict(ex=("ex_y", "ex_z"), ey=("ey_z", "ey_x"), ez=("ez_x", "ez_y"),
                 hx=("hx_y", "hx_z"), hy=("hy_z", "hy_x"), hz=("hz_x", "hz_y"))

    def inject(kind):
        for p in probes:
            comp = names[p["field"]]
            if (p["field"] < 3) != (kind == "e"):
                continue
            i = tuple(int(v) for v in p["idx"])
            s1, s2 = split[comp]
            v = p["values"]
            if "is_source" in p:
                f[s1][i] = f[s1][i] + v[n]
                f[s2][i] = f[s2][i] + v[n]
                f[comp][i] = f[s1][i] + f[s2][i]
            v[n] = f[comp][i]

    for n in range(Nt):
        hx, hy, hz = f["hx"], f["hy"], f["hz"]
        hyp = np.concatenate([hy, z(1, Ny + 1, Nz)], axis=0)  # zero plane beyond x = Nx
        hzp = np.concatenate([hz, z(1, Ny, Nz + 1)], axis=0)
        s = (slice(None), slice(1, Ny), slice(1, Nz))
        f["ex_y"][s] = c["Ca_ex_y"] * f["ex_y"][s] + c["Cb_ex_y"] * (hz[:, 1:Ny, 1:Nz] - hz[:, 0:Ny - 1, 1:Nz])
        f["ex_z"][s] = c["Ca_ex_z"] * f["ex_z"][s] + c["Cb_ex_z"] * (hy[:, 1:Ny, 1:Nz] - hy[:, 1:Ny, 0:Nz - 1])
        s = (slice(1, None), slice(None), slice(1, Nz))
        f["ey_z"][s] = c["Ca_ey_z"] * f["ey_z"][s] + c["Cb_ey_z"] * (hx[1:, :, 1:Nz] - hx[1:, :, 0:Nz - 1])
        f["ey_x"][s] = c["Ca_ey_x"] * f["ey_x"][s] + c["Cb_ey_x"] * (hzp[1:, :, 1:Nz] - hzp[:-1, :, 1:Nz])
        s = (slice(1, None), slice(1, Ny), slice(None))
        f["ez_x"][s] = c["Ca_ez_x"] * f["ez_x"][s] + c["Cb_ez_x"] * (hyp[1:, 1:Ny, :] - hyp[:-1, 1:Ny, :])
        f["ez_y"][s] = c["Ca_ez_y"] * f["ez_y"][s] + c["Cb_ez_y"] * (hx[1:, 1:Ny, :] - hx[1:, 0:Ny - 1, :])
        f["ex"][...] = f["ex_y"] + f["ex_z"]
        f["ey"][...] = f["ey_z"] + f["ey_x"]
        f["ez"][...] = f["ez_x"] + f["ez_y"]
        inject("e")
        ex, ey, ez = f["ex"], f["ey"], f["ez"]
        f["hx_y"][1:] = c["Da_hx_y"] * f["hx_y"][1:] + (c["Db_hx_y2"] * ez[1:, 1:, :] - c["Db_hx_y1"] * ez[1:, :-1, :])
        f["hx_z"][1:] = c["Da_hx_z"] * f["hx_z"][1:] + (c["Db_hx_z2"] * ey[1:, :, 1:] - c["Db_hx_z1"] * ey[1:, :, :-1])
        f["hy_z"][...] = c["Da_hy_z"] * f["hy_z"] + (c["Db_hy_z2"] * ex[:, :, 1:] - c["Db_hy_z1"] * ex[:, :, :-1])
        f["hy_x"][...] = c["Da_hy_x"] * f["hy_x"] + (c["Db_hy_x2"] * ez[1:] - c["Db_hy_x1"] * ez[:-1])
        f["hz_x"][...] = c["Da_hz_x"] * f["hz_x"] + (c["Db_hz_x2"] * ey[1:] - c["Db_hz_x1"] * ey[:-1])
        f["hz_y"][...] = c["Da_hz_y"] * f["hz_y"] + (c["Db_hz_y2"] * ex[:, 1:, :] - c["Db_hz_y1"] * ex[:, :-1, :])
        f["hx"][...] = f["hx_y"] + f["hx_z"]
        f["hy"][...] = f["hy_z"] + f["hy_x"]
        f["hz"][...] = f["hz_x"] + f["hz_y"]
        inject("h")
        for m in monitors:
            if n > 0 and n % m["n_step"]:
                continue
            comp = names[m["field"]]
            arr, pos, ax = f[comp], m["position"], m["axis"]
            # the CPU packing holds no x = 0 plane of Ey, Ez and Hx
            lead = 1 if comp in ("ey", "ez", "hx") else 0
            if ax == 0:
                plane = arr[pos]
            elif ax == 1:
                plane = arr[lead:, pos, :]
            else:
                plane = arr[lead:, :, pos]
            mn = n // m["n_step"]
            if "dtft_phase" in m:
                ph = m["dtft_phase"][mn]
                acc = m["values"]
                re = acc.real + plane[None] * ph.real[:, None, None]
                im = acc.imag + plane[None] * ph.imag[:, None, None]
                acc.real, acc.imag = re, im
            else:
                m["values"][mn] = plane
    return f
