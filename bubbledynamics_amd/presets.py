"""Parameter presets: the reference's own run and the benchmark workloads.

``reference_westervelt()`` re-derives the reference's host constants exactly
(Westervelt_relax_fft_Pressure_Energy.cu:356-417).  The workload presets
describe the 192^3 tone-burst run and the smaller grids used in tests.
"""
from __future__ import annotations

import math

from ._ffi import BD_BC_MUR, BD_BC_ZERO, BD_SRC_ARRAY, BD_SRC_CAP, WesterveltParams


def reference_westervelt() -> tuple[WesterveltParams, int, int]:
    """Reference run constants (W:356-417).  Returns (params, Nxy, Nz)."""
    p = WesterveltParams()
    fr = 1.253 * 10.0 ** 6                       # W:356
    d = 0.15                                     # W:357
    a = 0.151 / 2.0                              # W:358
    h = d - math.sqrt(d * d - a * a)             # W:359
    a_hole = 0.041 / 2.0                         # W:360
    p.fr, p.d = fr, d
    p.alpha = math.asin(a / d)                   # W:362
    p.alpha_hole = math.asin(a_hole / d)         # W:363
    p.rho, p.c_inf = 1000.0, 1500.0              # W:366-367
    p.abs_water = 0.026 * (fr / 10.0 ** 6) ** 2  # W:371
    p.delta = (2.0 * p.abs_water * p.c_inf ** 3) / (2.0 * math.pi * fr) ** 2  # W:372
    p.beta = 3.5                                 # W:373
    p.rho_tissue, p.c_inf_tissue = 1055.0, 1600.0  # W:376-377
    p.beta_tissue = 4.5                          # W:381
    p.abs_tissue = 7.2 * fr / 10.0 ** 6          # W:382
    p.delta_tissue = 4.2318 * 10.0 ** -5         # W:389
    p.c1, p.c2 = 4.03536, 6.14150                # W:390-391
    p.tau1, p.tau2 = 1.6178 * 10.0 ** -8, 1.1494 * 10.0 ** -7  # W:392-393
    eps, E = 0.89, 100.0                         # W:396-397
    u0 = math.sqrt(2.0 * E * eps / (p.rho * p.c_inf * 2.0 * math.pi * d * h))  # W:398
    p.p0 = p.rho * p.c_inf * u0                  # W:400
    p.water_tissue = 0.12                        # W:402
    p.xy_min, xy_max, p.dxy = 0.0, 0.085, 0.0002  # W:408-409
    nxy = 1 + round((xy_max - p.xy_min) / p.dxy)  # W:410
    z_max, p.z_min, p.dz = 0.2, -0.0002, 0.0001  # W:412-413
    nz = 1 + round((z_max - p.z_min) / p.dz)     # W:414
    p.dt = 0.01 / fr                             # W:416
    return p, nxy, nz


def reference_scaled(nxy: int = 48, nz: int = 64) -> WesterveltParams:
    """The reference's two-layer water/muscle physics on a small grid.

    Media, relaxation and the Mur boundaries are the reference's; the grid is
    coarsened and the cap shrunk so that a test grid holds source, interface
    and focus.  The interface sits at the middle plane.
    """
    p, _, _ = reference_westervelt()
    p.fr = 1.0e6
    p.dxy = p.dz = 1.5e-4
    p.z_min = -2 * p.dz
    p.xy_min = 0.0
    p.dt = 0.3 * p.dz / p.c_inf_tissue
    p.water_tissue = p.z_min + (nz // 2) * p.dz - 0.5 * p.dz
    p.d = 0.6 * (nz - 3) * p.dz
    p.alpha = math.asin(0.45)
    p.alpha_hole = math.asin(0.1)
    p.p0 = 1.0e6
    return p


def _homogeneous(p: WesterveltParams, c: float, rho: float, beta: float, delta: float,
                 relax: tuple[tuple[float, float], tuple[float, float]] | None) -> None:
    p.rho = p.rho_tissue = rho
    p.c_inf = p.c_inf_tissue = c
    p.beta = p.beta_tissue = beta
    p.delta = p.delta_tissue = delta
    if relax is None:
        # no relaxing region: the whole domain is the reference's "water" branch
        p.water_tissue = 1.0e9
        p.c1 = p.c2 = 0.0
        p.tau1 = p.tau2 = 1.0e-9
    else:
        (p.tau1, p.c1), (p.tau2, p.c2) = relax
        # every plane is in the relaxing ("tissue") branch
        p.water_tissue = -1.0e9


def tone_burst_192(n: int = 192) -> WesterveltParams:
    """Plain Westervelt 3-D workload: 192^3, dx=0.15 mm, CFL 0.3, water-like
    medium with two relaxation mechanisms everywhere, 1 MHz 8-cycle tone
    burst at 1 MPa from a spherical cap on the z=0 plane (iz=2), 20-cell
    sponge.  The grid (28.65 mm per side) cannot hold a 64 mm / 64 mm cap, so
    the cap keeps F-number 1 and is scaled to focal length 22.5 mm (focus at
    iz=152) and aperture 22.5 mm.  Quarter domain with mirror planes at x=0
    and y=0 (the reference's layout), Mur faces outside the sponge.
    """
    p = WesterveltParams()
    c, dx = 1500.0, 1.5e-4
    fr = 1.0e6
    alpha_w = 0.026 * (fr / 1e6) ** 2
    delta = 2.0 * alpha_w * c ** 3 / (2 * math.pi * fr) ** 2  # 4.445e-6 m^2/s
    _homogeneous(p, c, 1000.0, 3.5, delta, ((1.0e-9, 1.5), (1.0e-8, 3.0)))
    p.abs_water = p.abs_tissue = alpha_w
    p.fr = fr
    p.dxy = p.dz = dx
    p.xy_min = 0.0
    p.z_min = -2 * dx
    p.dt = 0.3 * dx / c
    p.d = 150 * dx * (n / 192.0)
    p.alpha = math.asin(0.5)       # aperture radius d/2: F-number 1
    p.alpha_hole = 0.0
    p.p0 = 1.0e6
    p.boundary = BD_BC_MUR
    p.source = BD_SRC_CAP
    p.sponge_cells = max(4, round(20 * n / 192))
    p.sponge_alpha = 0.015 * 20 / p.sponge_cells
    p.burst_cycles = 8.0
    rmax = p.d * math.tan(p.alpha)
    p.burst_t0 = p.d * (math.sqrt(1.0 + (rmax / p.d) ** 2) - 1.0) / p.c_inf
    return p


def tone_burst_focus(p: WesterveltParams, n: int) -> tuple[int, int, int]:
    iz = 2 + round(p.d / p.dz)
    return 0, 0, min(iz, n - 4)


def phased_array(n: int = 128, n_el: int = 8, dx: float = 2.0e-4) -> WesterveltParams:
    """Full-domain phased-array workload (128^3, dx=0.2 mm, 1 MHz, CFL 0.3):
    an 8x8 tiling of square piston elements on plane iz=2, boundaries held at
    zero behind a sponge.  Amplitudes and phases are per-element inputs."""
    p = WesterveltParams()
    c = 1500.0
    fr = 1.0e6
    alpha_w = 0.026
    delta = 2.0 * alpha_w * c ** 3 / (2 * math.pi * fr) ** 2
    _homogeneous(p, c, 1000.0, 3.5, delta, ((1.0e-9, 1.5), (1.0e-8, 3.0)))
    p.fr = fr
    p.dxy = p.dz = dx
    p.full_domain = 1
    p.xy_min = -(n - 1) / 2 * dx
    p.z_min = -2 * dx
    p.dt = 0.3 * dx / c
    p.boundary = BD_BC_ZERO
    p.source = BD_SRC_ARRAY
    p.sponge_cells = max(4, round(12 * n / 128))
    p.sponge_alpha = 0.02 * 12 / p.sponge_cells
    cells = max(1, round(0.6 * (n - 2 * p.sponge_cells) / n_el))
    p.arr_nx = p.arr_ny = n_el
    p.arr_pitch = cells * dx
    # element edges fall half-way between grid nodes
    p.arr_x0 = p.arr_y0 = -n_el * p.arr_pitch / 2 + (0.5 * dx if n % 2 else 0.0)
    p.p0 = 1.0e6
    return p
