"""Cylindrical (r, theta, z) finite-volume grid (upstream core/grid.py:26-108).

Cell (i, j, k) spans r_f[i..i+1], th_f[j..j+1], z_f[k..k+1]; z = 0 is the top (feed) end.
The per-ring geometric factors are what the device model stores (V, Az, Ar, Ath, drc, arc).
"""
from __future__ import annotations

import numpy as np


class Grid:
    def __init__(self, R, L, Nr, Nth, Nz, wall_thickness=0.002, radial_spacing="uniform"):
        self.R, self.L, self.Nr, self.Nth, self.Nz = R, L, Nr, Nth, Nz
        self.wall_thickness = wall_thickness
        self.radial_spacing = radial_spacing
        self.r_f = R * np.sqrt(np.linspace(0.0, 1.0, Nr + 1)) if radial_spacing == "equal_area" \
            else np.linspace(0.0, R, Nr + 1)
        self.r_c = 0.5 * (self.r_f[1:] + self.r_f[:-1])
        self.dr = np.diff(self.r_f)
        self.dth = 2.0 * np.pi / Nth
        self.th_f = np.linspace(0.0, 2.0 * np.pi, Nth + 1)
        self.th_c = 0.5 * (self.th_f[1:] + self.th_f[:-1])
        self.dz = L / Nz
        self.z_f = np.linspace(0.0, L, Nz + 1)
        self.z_c = 0.5 * (self.z_f[1:] + self.z_f[:-1])
        ring = 0.5 * (self.r_f[1:] ** 2 - self.r_f[:-1] ** 2)
        self.Az = (ring * self.dth)[:, None]
        self.V = (ring * self.dth * self.dz)[:, None, None]
        self.Ar = (self.r_f[1:-1] * self.dth * self.dz)[:, None, None]
        self.Ath = (self.dr * self.dz)[:, None, None]
        self.drc = np.diff(self.r_c)[:, None, None]
        self.arc = (self.r_c * self.dth)[:, None, None]
        self.dzc = self.dz
        self.A_wall_in = R * self.dth * self.dz
        Ro = R + wall_thickness
        self.A_wall_out = Ro * self.dth * self.dz
        self.V_wall = 0.5 * (Ro ** 2 - R ** 2) * self.dth * self.dz
        self.A_wall_axial = 0.5 * (Ro ** 2 - R ** 2) * self.dth
        self.R_wall_mid = R + 0.5 * wall_thickness
        self.half_dr_outer = R - self.r_c[-1]
        self.cross_section = np.pi * R ** 2
        self.bed_volume = self.cross_section * L
        self.ncell = Nr * Nth * Nz
        self.shape = (Nr, Nth, Nz)
        self.wall_shape = (Nth, Nz)

    @classmethod
    def from_config(cls, cfg) -> "Grid":
        return cls(R=cfg.geometry.bed_radius, L=cfg.geometry.bed_length, Nr=cfg.grid.Nr, Nth=cfg.grid.Ntheta,
                   Nz=cfg.grid.Nz, wall_thickness=cfg.geometry.wall_thickness,
                   radial_spacing=cfg.grid.radial_spacing)

    def cell_centres(self):
        return np.meshgrid(self.r_c, self.th_c, self.z_c, indexing="ij")

    def volume_full(self):
        return np.broadcast_to(self.V, self.shape)

    def integrate(self, field3d) -> float:
        return float(np.sum(field3d * self.V))

    def coverage_fraction(self, theta_deg, z_frac):
        t0, t1 = np.radians(theta_deg[0]), np.radians(theta_deg[1])
        z0, z1 = z_frac[0] * self.L, z_frac[1] * self.L
        ov_t = np.clip(np.minimum(self.th_f[1:], t1) - np.maximum(self.th_f[:-1], t0), 0, None) / self.dth
        ov_z = np.clip(np.minimum(self.z_f[1:], z1) - np.maximum(self.z_f[:-1], z0), 0, None) / self.dz
        return ov_t[:, None] * ov_z[None, :]

    def describe(self) -> str:
        return (f"Grid: Nr={self.Nr} Ntheta={self.Nth} Nz={self.Nz} ({self.ncell} cells), "
                f"R={self.R} m, L={self.L} m, dr={self.dr.mean():.3e} m, dz={self.dz:.3e} m")
