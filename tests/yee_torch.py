"""Unrolled, differentiable Yee time stepping in torch (test helper).

The same split-field equations as tests/yee_numpy.py (global component arrays, CPU-packed
coefficients of rfnetwork/solver.py:1463-1523 with gpu=False, E: Ca*s + Cb*(h+ - h),
H: Da*s + (Db2*e - Db1*e-), every probe injected into both splits and then recorded as the total,
rfnetwork/core/src/solver.cpp:986-994,1149-1158), written out of place so that torch autograd can
run the whole time loop backwards.  It is the independent reference for the engine's hand-written
adjoint: the gradient of sum_pn g[p, n] y_p[n] with respect to the six E-split Cb arrays.
"""
import numpy as np
import torch
import torch.nn.functional as F

E_CB = ("Cb_ex_y", "Cb_ex_z", "Cb_ey_z", "Cb_ey_x", "Cb_ez_x", "Cb_ez_y")
PEC_X = ("Cb_ey_z", "Cb_ey_x", "Ca_ey_z", "Ca_ey_x", "Cb_ez_x", "Cb_ez_y", "Ca_ez_x", "Ca_ez_y",
         "Db_hx_y1", "Db_hx_y2", "Db_hx_z1", "Db_hx_z2", "Da_hx_y", "Da_hx_z")


def run(coefficients, probes, Nx, Ny, Nz, Nt, leaves=E_CB):
    """Returns (series [P, Nt] tensor, {name: leaf tensor}) in float64; probes as packed for the
    CPU solver (global Yee indices, "is_source" key present for injected probes)."""
    dt = torch.float64
    c, leaf = {}, {}
    for k, v in coefficients.items():
        t = torch.tensor(np.asarray(v, np.float64), dtype=dt)
        if k in leaves:
            t.requires_grad_(True)
            leaf[k] = t
        c[k] = t
    keep = torch.ones(Nx, 1, 1, dtype=dt)
    keep[Nx - 1] = 0  # the PEC wall at x = Nx zeroes these coefficient planes
    for k in PEC_X:
        c[k] = c[k] * keep
    z = lambda *s: torch.zeros(s, dtype=dt)
    f = dict(ex_y=z(Nx, Ny + 1, Nz + 1), ex_z=z(Nx, Ny + 1, Nz + 1), ey_z=z(Nx + 1, Ny, Nz + 1),
             ey_x=z(Nx + 1, Ny, Nz + 1), ez_x=z(Nx + 1, Ny + 1, Nz), ez_y=z(Nx + 1, Ny + 1, Nz),
             hx_y=z(Nx + 1, Ny, Nz), hx_z=z(Nx + 1, Ny, Nz), hy_z=z(Nx, Ny + 1, Nz), hy_x=z(Nx, Ny + 1, Nz),
             hz_x=z(Nx, Ny, Nz + 1), hz_y=z(Nx, Ny, Nz + 1))
    names = ("ex", "ey", "ez", "hx", "hy", "hz")
    split = dict(ex=("ex_y", "ex_z"), ey=("ey_z", "ey_x"), ez=("ez_x", "ez_y"),
                 hx=("hx_y", "hx_z"), hy=("hy_z", "hy_x"), hz=("hz_x", "hz_y"))
    src = [torch.tensor(np.asarray(p["values"][:Nt], np.float64), dtype=dt) for p in probes]
    rec = [[None] * Nt for _ in probes]

    def total(comp):
        a, b = split[comp]
        return f[a] + f[b]

    def inject(kind, n):
        for k, p in enumerate(probes):
            if (p["field"] < 3) != (kind == "e"):
                continue
            comp = names[p["field"]]
            i = tuple(torch.tensor([int(v)]) for v in p["idx"])
            s1, s2 = split[comp]
            if "is_source" in p:
                v = src[k][n].reshape(1)
                f[s1] = f[s1].index_put(i, f[s1][i] + v)
                f[s2] = f[s2].index_put(i, f[s2][i] + v)
            rec[k][n] = (f[s1][i] + f[s2][i]).reshape(())

    for n in range(Nt):
        hx, hy, hz = total("hx"), total("hy"), total("hz")
        hyp = torch.cat([hy, z(1, Ny + 1, Nz)], 0)  # zero plane beyond x = Nx
        hzp = torch.cat([hz, z(1, Ny, Nz + 1)], 0)
        g = f
        g["ex_y"] = F.pad(c["Ca_ex_y"] * g["ex_y"][:, 1:Ny, 1:Nz]
                          + c["Cb_ex_y"] * (hz[:, 1:Ny, 1:Nz] - hz[:, 0:Ny - 1, 1:Nz]), (1, 1, 1, 1))
        g["ex_z"] = F.pad(c["Ca_ex_z"] * g["ex_z"][:, 1:Ny, 1:Nz]
                          + c["Cb_ex_z"] * (hy[:, 1:Ny, 1:Nz] - hy[:, 1:Ny, 0:Nz - 1]), (1, 1, 1, 1))
        g["ey_z"] = F.pad(c["Ca_ey_z"] * g["ey_z"][1:, :, 1:Nz]
                          + c["Cb_ey_z"] * (hx[1:, :, 1:Nz] - hx[1:, :, 0:Nz - 1]), (1, 1, 0, 0, 1, 0))
        g["ey_x"] = F.pad(c["Ca_ey_x"] * g["ey_x"][1:, :, 1:Nz]
                          + c["Cb_ey_x"] * (hzp[1:, :, 1:Nz] - hzp[:-1, :, 1:Nz]), (1, 1, 0, 0, 1, 0))
        g["ez_x"] = F.pad(c["Ca_ez_x"] * g["ez_x"][1:, 1:Ny, :]
                          + c["Cb_ez_x"] * (hyp[1:, 1:Ny, :] - hyp[:-1, 1:Ny, :]), (0, 0, 1, 1, 1, 0))
        g["ez_y"] = F.pad(c["Ca_ez_y"] * g["ez_y"][1:, 1:Ny, :]
                          + c["Cb_ez_y"] * (hx[1:, 1:Ny, :] - hx[1:, 0:Ny - 1, :]), (0, 0, 1, 1, 1, 0))
        inject("e", n)
        ex, ey, ez = total("ex"), total("ey"), total("ez")
        g["hx_y"] = F.pad(c["Da_hx_y"] * g["hx_y"][1:] + (c["Db_hx_y2"] * ez[1:, 1:, :] - c["Db_hx_y1"] * ez[1:, :-1, :]),
                          (0, 0, 0, 0, 1, 0))
        g["hx_z"] = F.pad(c["Da_hx_z"] * g["hx_z"][1:] + (c["Db_hx_z2"] * ey[1:, :, 1:] - c["Db_hx_z1"] * ey[1:, :, :-1]),
                          (0, 0, 0, 0, 1, 0))
        g["hy_z"] = c["Da_hy_z"] * g["hy_z"] + (c["Db_hy_z2"] * ex[:, :, 1:] - c["Db_hy_z1"] * ex[:, :, :-1])
        g["hy_x"] = c["Da_hy_x"] * g["hy_x"] + (c["Db_hy_x2"] * ez[1:] - c["Db_hy_x1"] * ez[:-1])
        g["hz_x"] = c["Da_hz_x"] * g["hz_x"] + (c["Db_hz_x2"] * ey[1:] - c["Db_hz_x1"] * ey[:-1])
        g["hz_y"] = c["Da_hz_y"] * g["hz_y"] + (c["Db_hz_y2"] * ex[:, 1:, :] - c["Db_hz_y1"] * ex[:, :-1, :])
        inject("h", n)
    series = torch.stack([torch.stack(r) for r in rec])
    return series, leaf
