"""Writers for keyword decks and mesh files in the reference's input grammar.

The grammar is the one parsed by the reference's ``Mesh::getDomainInfo``
(src/sources/Mesh.cpp:39-435) and ``Mesh::readMeshFile`` (Mesh.cpp:475-550):
whole-line ``*KEYWORD`` markers, ``$`` comment lines, whitespace-separated
values; no blank lines (the reference throws on one, Mesh.cpp:50).

Every deck written here carries a zero-power ``*VOLUMETRIC_HEAT_SOURCE``:
without one the reference indexes an empty heat-source table
(src/sources/CFDSolverManager.cpp:1707-1716).

The named cases are the workloads the solver is measured and checked on:

* ``tiny``      22x14x14 nodes, 40 steps: full-field fixtures every step.
* ``tiny_cond`` same grid, upwind, ``*FLUID_FLOW_OFF``: conduction-only path.
* ``tiny_evap`` same grid, evaporation loss and the Fe-O surfactant
                surface-tension model (a_oxygen > 0).
* ``tiny_vap``  same grid, evaporation loss only (constant dgamma/dT).
* ``tiny_vol``  same grid, Gaussian laser plus a 100 W volumetric source.
* ``tiny_out``  same grid, 20 steps with the laser on throughout, output every
                5 steps (VTS/PVD, residual log, final geometry).
* ``tiny_tool`` same grid, a two-track toolpath file with a laser-off hatch.
* ``tiny_powder`` powder-bed mode, two layers; the toolpath's off->on
                transition adds the second layer mid-run.
* ``tiny_move``  24x16x14 nodes, ``*ACTIVATE_MOVING_MESH 2 2``: the middle x
                and y zones follow the beam (+x, +y hatch, -x).
* ``tiny_move_powder`` the moving mesh over a two-layer powder bed (the
                configuration of the reference's multi-layer examples).
* ``tiny_species`` the tiny grid with species transport (*SOLVE_SPECIES) and
                the G/R post-processing (*CALCULATE_SOLIDIFICATION_PARAMETERS).
* ``tiny_move_species`` the moving mesh with species and G/R (the remap then
                carries the concentration fields too).
* ``tiny_fs``   the tiny_out outputs plus the energy-based free surface (at
                every output) and an extracted section every 20 us.
* ``tiny_dtvar`` a *CONTROL_TIMESTEP_VARIABLE table: dt 5, 2.5, then 4 us.
* ``tiny_heattable`` a *VOLUMETRIC_HEAT_SOURCE_TABLE whose row changes with time.
* ``tiny_steady`` dt = 100 s: the reference's steady-state mode, two steps of
                up to 150 iterations with the beam fixed at xstart, conduction
                only (with the flow on, the reference's steady run turns NaN
                on this grid).
* ``plate120``  120x48x32 cells (2.0x0.6x0.4 mm), 400 steps, P=200 W,
                r=50 um, 1 m/s: the CPU-runnable single-track case.
* ``track400``  400x160x80 cells (4.0x1.2x0.6 mm, ~5.1M cells), P=300 W,
                r=60 um, 0.8 m/s, dt=2e-6 s: the single-GPU throughput case.
* ``multilayer`` two 40 um powder layers (conductivity 0.3x bulk) of four
                tracks each, 80 um hatch, alternating +x/-x, 3.0x1.0x0.8 mm
                on 480x160x126 cells (~9.7M), P=250 W, 1 m/s, dt=2e-6 s; the
                laser stays on across the hatch jumps inside a layer, so only
                the layer change (laser off, then on) adds the second layer.
* ``multilayer_small`` the same toolpath pattern (two layers, four tracks,
                80 um hatch, alternating scan) on a 60x48x24-cell plate that
                the unmodified reference runs (dt=5e-6 s, 100 um tracks).
"""
import os

# Synthetic alloy (rho 4000, cp 700, k 25, Ts 1878, Tl 1928, L 2.9e5, mu 5e-3).
ACPA = 1.0e-6
ACPB = 700.0
TS = 1878.0
TL = 1928.0
LATENT = 2.9e5
HSMELT = 0.5 * ACPA * TS * TS + ACPB * TS
HLCAL = HSMELT + 0.5 * (ACPA * TS + ACPB + ACPB) * (TL - TS)
HLFRIZ = HLCAL + LATENT


def _fmt(v):
    return repr(float(v)) if not isinstance(v, str) else v


def write_mesh(path, x, y, z):
    """x/y/z: (lengths, ncv, exponents) per axis, one entry per zone."""
    lines = []
    for tag, (lens, ncv, expo) in (("X", x), ("Y", y), ("Z", z)):
        lines.append("*%s_GRID_PARAMETERS" % tag)
        lines.append("$ number of zones")
        lines.append("  %d" % len(lens))
        lines.append("$ zone lengths (m)")
        lines.append("  " + " ".join(_fmt(v) for v in lens))
        lines.append("$ control volumes per zone")
        lines.append("  " + " ".join(str(int(v)) for v in ncv))
        lines.append("$ exponents")
        lines.append("  " + " ".join(_fmt(v) for v in expo))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_deck(path, mesh_path, *, keyword="case", finaltime, maxit=100, tol=1e-6,
               dt, lasertime, xstart, ystart, speed, power, eta=0.3, radius,
               factor=2.0, discretization="power_law", flow=True, urf=(0.8,) * 5,
               tamb=300.0, tpre=300.0, htc=(10.0, 10.0, 10.0, 10.0), emiss=0.4,
               momentum=(5e-5, -2.6e-4, 1e7, -1.0), material=None, evaporation=None,
               volumetric=None, outtime=None, scalar_out=("temp",), final_geometry=None,
               toolpath=None, powder=None, moving=None, species=None, solidification=False,
               free_surface=None, section=None, dt_table=None, heat_table=None, extra=()):
    """toolpath: rows of (t, x, y, z, laser_on, track) written to <deck>_tool.crs;
    powder: the 8 *POWDER_BED_MODE values; moving: the two 1-based zone
    indices of *ACTIVATE_MOVING_MESH; species: the 4 *SOLVE_SPECIES values
    (initial concentration, mass diffusivity, partition coefficient,
    relaxation); solidification: add *CALCULATE_SOLIDIFICATION_PARAMETERS;
    free_surface: the 8 *TRACK_ENERGY_BASED_FREE_SURFACE values; section: the
    7 *OUTPUT_SECTION values (box min, box max, interval); dt_table: rows of
    (t_until, dt, outtime) for *CONTROL_TIMESTEP_VARIABLE (<deck>_dt.k);
    heat_table: rows of (t, P, eta, thickness, radius, factor) for
    *VOLUMETRIC_HEAT_SOURCE_TABLE (<deck>_heat.k), replacing the fixed source."""
    stem = os.path.splitext(os.path.abspath(path))[0]

    def table(suffix, rows):
        aux = stem + suffix
        with open(aux, "w") as f:
            for row in rows:
                f.write(" ".join(_fmt(v) for v in row) + "\n")
        return aux
    if material is None:
        material = (4000.0, 4000.0, 5e-3, TS, TL, HSMELT, HLFRIZ, ACPA, ACPB, ACPB,
                    0.0, 0.0, 25.0, 0.0, 25.0)
    L = []
    L += ["*KEYWORD_ID", keyword, "*DIRECTORY", keyword]
    L += ["*PARAMETERS", "accelGravity 9.8", "radBoltzmann 5.67e-8",
          "ambientTemp %s" % _fmt(tamb), "ambientPres 101.0e3", "gasConstant 8.314"]
    for name in scalar_out:
        L += ["*SCALAR_OUT", name]
    L += ["*CONTROL_TERMINATION", "$ total_time  max_iter  nonlin_tolerance",
          "  %s %d %s" % (_fmt(finaltime), maxit, _fmt(tol))]
    L += ["*CONTROL_TIMESTEP", "$ sim_timestep  output_timestep",
          "  %s %s" % (_fmt(dt), _fmt(finaltime if outtime is None else outtime))]
    if dt_table is not None:
        L += ["*CONTROL_TIMESTEP_VARIABLE", table("_dt.k", dt_table)]
    L += ["*SINGLE_TRACK_TOOL", "$ lasertime  xstart  ystart  vel",
          "  %s %s %s %s" % tuple(_fmt(v) for v in (lasertime, xstart, ystart, speed))]
    L += ["*GAUSS_LASER_PARAMETERS", "$ power  absorb  radius  factor",
          "  %s %s %s %s" % tuple(_fmt(v) for v in (power, eta, radius, factor))]
    if heat_table is not None:
        L += ["*VOLUMETRIC_HEAT_SOURCE_TABLE", table("_heat.k", heat_table)]
    else:
        L += ["*VOLUMETRIC_HEAT_SOURCE", "$ power apowseta heatthick heatrb factor",
              "  0.0 0.0 1e-5 5e-5 2.0" if volumetric is None else
              "  " + " ".join(_fmt(v) for v in volumetric)]
    m = material
    L += ["*MATERIAL_PROPERTIES", "$ dens denl viscos tsolid tliquid hsmelt hlfriz",
          "  " + " ".join(_fmt(v) for v in m[0:7]), "$ acpa acpb acpl",
          "  " + " ".join(_fmt(v) for v in m[7:10]), "$ thconsa thconsb thconsc thconla thconlb",
          "  " + " ".join(_fmt(v) for v in m[10:15])]
    L += ["*UNDER_RELAXATION", "  " + " ".join(_fmt(v) for v in urf)]
    L += ["*BOUNDARY_TEMPERATURES", "  " + " ".join(_fmt(tpre) for _ in range(5))]
    L += ["*BOUNDARY_ENERGY_EQUATION", "  " + " ".join(_fmt(v) for v in htc) + " " + _fmt(emiss)]
    L += ["*BOUNDARY_MOMENTUM_EQUATION", "  " + " ".join(_fmt(v) for v in momentum)]
    if discretization:
        L += ["*SELECT_DISCRETIZATION", discretization]
    if not flow:
        L += ["*FLUID_FLOW_OFF"]
    if evaporation is not None:
        L += ["*BOUNDARY_EVAPORATION", "  " + " ".join(_fmt(v) for v in evaporation)]
    if final_geometry is not None:
        L += ["*OUTPUT_FINAL_GEOMETRY", final_geometry]
    if powder is not None:
        L += ["*POWDER_BED_MODE", "$ numlayers ncvlayer layerheight density cpa cpb thcona thconb",
              "  " + " ".join(_fmt(v) for v in powder)]
    if moving is not None:
        L += ["*ACTIVATE_MOVING_MESH", "  %d %d" % tuple(moving)]
    if species is not None:
        L += ["*SOLVE_SPECIES", "$ initconc massdiff kp urfc", "  " + " ".join(_fmt(v) for v in species)]
    if solidification:
        L += ["*CALCULATE_SOLIDIFICATION_PARAMETERS"]
    if free_surface is not None:
        L += ["*TRACK_ENERGY_BASED_FREE_SURFACE",
              "$ lambda1 lambda2 surftens maxiter_energy tol_energy maxiter_bisect tol_bisect urf",
              "  " + " ".join(_fmt(v) for v in free_surface)]
    if section is not None:
        L += ["*OUTPUT_SECTION", "$ xmin ymin zmin xmax ymax zmax interval",
              "  " + " ".join(_fmt(v) for v in section)]
    L += list(extra)
    L += ["*MESH_FILE", mesh_path]
    if toolpath is not None:
        tool_path = stem + "_tool.crs"
        with open(tool_path, "w") as f:
            for row in toolpath:
                f.write(" ".join(_fmt(v) for v in row[:4]) + " %d %d\n" % (row[4], row[5]))
        L += ["*TOOLPATH_FILE", tool_path]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


# Toolpaths, rows (t, x, y, z, laser_on, track); a row's laser state holds on
# the interval that ends at it (BoundCondManager.cpp:216-235).
ZTOP = 0.25e-3
# two +x/-x passes joined by a laser-off hatch jump in +y
TOOL_TWO_TRACKS = [(0.0, 0.15e-3, 0.12e-3, ZTOP, 1, 1), (60e-6, 0.25e-3, 0.12e-3, ZTOP, 1, 1),
                   (70e-6, 0.25e-3, 0.18e-3, ZTOP, 0, 1), (130e-6, 0.17e-3, 0.18e-3, ZTOP, 1, 2),
                   (200e-6, 0.17e-3, 0.18e-3, ZTOP, 0, 2)]
# the same, with the laser-off -> on transition that adds a powder layer
TOOL_TWO_LAYERS = [(0.0, 0.15e-3, 0.15e-3, ZTOP, 1, 1), (60e-6, 0.25e-3, 0.15e-3, ZTOP, 1, 1),
                   (75e-6, 0.15e-3, 0.15e-3, ZTOP + 2e-5, 0, 1), (135e-6, 0.25e-3, 0.15e-3, ZTOP + 2e-5, 1, 2),
                   (200e-6, 0.25e-3, 0.15e-3, ZTOP + 2e-5, 0, 2)]
# moving-mesh path: +x at 1 m/s, a +y hatch, then -x (the remap's backward recurrence)
TOOL_MOVING = [(0.0, 0.25e-3, 0.15e-3, ZTOP, 1, 1), (60e-6, 0.31e-3, 0.15e-3, ZTOP, 1, 1),
               (70e-6, 0.31e-3, 0.19e-3, ZTOP, 0, 1), (130e-6, 0.25e-3, 0.19e-3, ZTOP, 1, 2),
               (200e-6, 0.25e-3, 0.19e-3, ZTOP, 0, 2)]
TOOL_MOVING_LAYERS = [(0.0, 0.25e-3, 0.15e-3, ZTOP, 1, 1), (60e-6, 0.31e-3, 0.15e-3, ZTOP, 1, 1),
                      (75e-6, 0.25e-3, 0.13e-3, ZTOP + 2e-5, 0, 1),
                      (135e-6, 0.31e-3, 0.13e-3, ZTOP + 2e-5, 1, 2),
                      (200e-6, 0.31e-3, 0.13e-3, ZTOP + 2e-5, 0, 2)]


def multilayer_toolpath(x0, x1, y0, hatch, ntracks, z0, layer, speed, hatch_time, change_time,
                        layers=2, t0=0.0):
    """Rows (t, x, y, z, laser_on, track) of a multi-track, multi-layer scan:
    per layer ntracks tracks between x0 and x1 (alternating +x/-x), joined by
    hatch jumps of `hatch` in +y taken with the laser on; between layers the
    laser is off while the head returns to (x0, y0) one layer higher.  Each
    laser off -> on transition adds one powder layer (BoundCondManager.cpp:
    265-273), so only the layer changes do."""
    rows = []
    t = t0
    track = 1
    for lay in range(layers):
        z = z0 + lay * layer
        x, y = x0, y0
        if lay == 0:
            rows.append((t, x, y, z, 1, track))
        else:
            t += change_time
            rows.append((t, x, y, z, 0, track - 1))
        for n in range(ntracks):
            if n > 0:
                y += hatch
                t += hatch_time
                rows.append((t, x, y, z, 1, track))
            x = x1 if x == x0 else x0
            t += abs(x1 - x0) / speed
            rows.append((t, x, y, z, 1, track))
            track += 1
    rows.append((t + hatch_time, rows[-1][1], rows[-1][2], rows[-1][3], 0, track - 1))
    return rows


# configs: two 40 um layers, 4 tracks each, 80 um hatch, 1 m/s
TOOL_MULTILAYER = multilayer_toolpath(0.5e-3, 2.5e-3, 0.38e-3, 80e-6, 4, 0.72e-3, 40e-6, 1.0, 80e-6, 0.5e-3)
TOOL_MULTILAYER_SMALL = multilayer_toolpath(0.3e-3, 0.4e-3, 0.18e-3, 80e-6, 4, 0.16e-3, 40e-6, 1.0,
                                            20e-6, 25e-6)
# powder: 2 layers, ncv per layer, 40 um, density, cpa, cpb, conductivity a, b (0.3 x 25 W/mK)
POWDER_MULTILAYER = (2, 6, 40e-6, 4000.0, 1.0e-6, 700.0, 0.0, 7.5)
POWDER_MULTILAYER_SMALL = (2, 4, 40e-6, 4000.0, 1.0e-6, 700.0, 0.0, 7.5)

CASES = {
    "tiny": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.2e-3 - 1e-9, maxit=100, dt=5e-6, lasertime=0.15e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=50e-6)),
    # x-lines of 1024 interior cells, the longest the line solvers take
    "tiny_long": dict(
        mesh=(([0.1e-3, 2.0e-3, 0.1e-3], [10, 1004, 10], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [3, 3], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [2, 4], [1.0, 1.0])),
        deck=dict(finaltime=0.03e-3 - 1e-9, maxit=40, dt=5e-6, lasertime=0.02e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=50e-6)),
    "tiny_cond": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.1e-3 - 1e-9, maxit=40, dt=5e-6, lasertime=0.08e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=50e-6,
                  discretization=None, flow=False)),
    "tiny_evap": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.1e-3 - 1e-9, maxit=60, dt=5e-6, lasertime=0.08e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=250.0, radius=50e-6,
                  momentum=(5e-5, -2.6e-4, 1e7, 0.02),
                  evaporation=(6.34e6, 3133.0, 0.001, 0.0558))),
    "tiny_vap": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.1e-3 - 1e-9, maxit=60, dt=5e-6, lasertime=0.08e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=250.0, radius=50e-6,
                  evaporation=(6.34e6, 3133.0, 0.001, 0.0558))),
    "tiny_vol": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.1e-3 - 1e-9, maxit=60, dt=5e-6, lasertime=0.08e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=150.0, radius=50e-6,
                  volumetric=(100.0, 0.3, 5e-5, 5e-5, 2.0))),
    "tiny_out": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.1e-3 - 1e-9, maxit=100, dt=5e-6, lasertime=0.15e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=50e-6,
                  outtime=2.5e-5, scalar_out=("temp", "uvel", "vvel", "wvel", "fracl", "enthalpy"),
                  final_geometry="tiny_out_geometry.txt")),
    "tiny_tool": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.15e-3 - 1e-9, maxit=60, dt=5e-6, lasertime=0.15e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=50e-6,
                  toolpath=TOOL_TWO_TRACKS)),
    "tiny_powder": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 6], [1.0, 1.0])),
        deck=dict(finaltime=0.15e-3 - 1e-9, maxit=60, dt=5e-6, lasertime=0.15e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=50e-6,
                  powder=(2, 2, 2.0e-5, 2400.0, 1.0e-6, 700.0, 0.0, 7.5),
                  toolpath=TOOL_TWO_LAYERS)),
    "tiny_move": dict(
        mesh=(([0.15e-3, 0.2e-3, 0.15e-3], [6, 10, 6], [1.0, 1.0, 1.0]),
              ([0.1e-3, 0.1e-3, 0.1e-3], [4, 6, 4], [1.0, 1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.15e-3 - 1e-9, maxit=60, dt=5e-6, lasertime=0.15e-3,
                  xstart=0.25e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=40e-6,
                  toolpath=TOOL_MOVING, moving=(2, 2))),
    "tiny_move_powder": dict(
        mesh=(([0.15e-3, 0.2e-3, 0.15e-3], [6, 10, 6], [1.0, 1.0, 1.0]),
              ([0.1e-3, 0.1e-3, 0.1e-3], [4, 6, 4], [1.0, 1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 6], [1.0, 1.0])),
        deck=dict(finaltime=0.15e-3 - 1e-9, maxit=60, dt=5e-6, lasertime=0.15e-3,
                  xstart=0.25e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=40e-6,
                  powder=(2, 2, 2.0e-5, 2400.0, 1.0e-6, 700.0, 0.0, 7.5),
                  toolpath=TOOL_MOVING_LAYERS, moving=(2, 2))),
    "tiny_species": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.15e-3 - 1e-9, maxit=60, dt=5e-6, lasertime=0.1e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=50e-6,
                  species=(1.0, 5e-9, 0.5, 0.7), solidification=True)),
    "tiny_move_species": dict(
        mesh=(([0.15e-3, 0.2e-3, 0.15e-3], [6, 10, 6], [1.0, 1.0, 1.0]),
              ([0.1e-3, 0.1e-3, 0.1e-3], [4, 6, 4], [1.0, 1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.15e-3 - 1e-9, maxit=60, dt=5e-6, lasertime=0.15e-3,
                  xstart=0.25e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=40e-6,
                  toolpath=TOOL_MOVING, moving=(2, 2), species=(1.0, 5e-9, 0.5, 0.7),
                  solidification=True)),
    "tiny_fs": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.1e-3 - 1e-9, maxit=100, dt=5e-6, lasertime=0.08e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=50e-6,
                  outtime=2.5e-5, scalar_out=("temp", "tempr", "zr", "fracl"),
                  free_surface=(-1.0e6, 1.0e6, 1.8, 20, 1e-6, 30, 1e-16, 0.5),
                  section=(0.2e-3, 0.1e-3, 0.15e-3, 0.3e-3, 0.15e-3, 0.25e-3, 2e-5))),
    "tiny_dtvar": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.12e-3, maxit=60, dt=5e-6, lasertime=0.1e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=200.0, radius=50e-6,
                  dt_table=[(3e-5, 5e-6, 1e-4), (6e-5, 2.5e-6, 1e-4), (1.0, 4e-6, 1e-4)])),
    "tiny_heattable": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=0.1e-3 - 1e-9, maxit=60, dt=5e-6, lasertime=0.1e-3,
                  xstart=0.15e-3, ystart=0.15e-3, speed=1.0, power=120.0, radius=50e-6,
                  heat_table=[(2e-5, 60.0, 0.3, 5e-5, 5e-5, 2.0), (5e-5, 120.0, 0.35, 6e-5, 4e-5, 2.0),
                              (1.0, 80.0, 0.3, 4e-5, 5e-5, 2.0)])),
    "tiny_steady": dict(
        mesh=(([0.1e-3, 0.3e-3, 0.1e-3], [4, 12, 4], [-1.2, 1.0, 1.2]),
              ([0.15e-3, 0.15e-3], [6, 6], [1.0, 1.0]),
              ([0.15e-3, 0.1e-3], [4, 8], [1.0, 1.0])),
        deck=dict(finaltime=200.0, maxit=150, dt=100.0, lasertime=0.1e-3,
                  xstart=0.25e-3, ystart=0.15e-3, speed=1.0, power=120.0, radius=50e-6, flow=False)),
    "plate120": dict(
        mesh=(([0.5e-3, 1.0e-3, 0.5e-3], [20, 80, 20], [-1.2, 1.0, 1.2]),
              ([0.2e-3, 0.2e-3, 0.2e-3], [16, 16, 16], [1.0, 1.0, 1.0]),
              ([0.2e-3, 0.2e-3], [16, 16], [1.0, 1.0])),
        deck=dict(finaltime=2.0e-3, maxit=100, dt=5e-6, lasertime=1.0e-3,
                  xstart=0.5e-3, ystart=0.3e-3, speed=1.0, power=200.0, radius=50e-6)),
    "track400": dict(
        mesh=(([4.0e-3], [400], [1.0]), ([1.2e-3], [160], [1.0]), ([0.6e-3], [80], [1.0])),
        deck=dict(finaltime=3.999e-3, maxit=100, dt=2e-6, lasertime=3.0e-3,
                  xstart=0.4e-3, ystart=0.6e-3, speed=0.8, power=300.0, radius=60e-6)),
    # z: a 0.64 mm base zone and a 0.08 mm top zone that the powder layers
    # extend by 2 x 40 um (6 cells each): 126 cells, 0.8 mm
    "multilayer": dict(
        mesh=(([3.0e-3], [480], [1.0]), ([1.0e-3], [160], [1.0]),
              ([0.64e-3, 0.08e-3], [102, 12], [1.0, 1.0])),
        deck=dict(finaltime=TOOL_MULTILAYER[-1][0], maxit=100, dt=2e-6, lasertime=TOOL_MULTILAYER[-1][0],
                  xstart=0.5e-3, ystart=0.38e-3, speed=1.0, power=250.0, radius=50e-6,
                  powder=POWDER_MULTILAYER, toolpath=TOOL_MULTILAYER)),
    "multilayer_small": dict(
        mesh=(([0.75e-3], [60], [1.0]), ([0.6e-3], [48], [1.0]),
              ([0.12e-3, 0.04e-3], [12, 4], [1.0, 1.0])),
        deck=dict(finaltime=TOOL_MULTILAYER_SMALL[-1][0] - 1e-9, maxit=40, dt=5e-6,
                  lasertime=TOOL_MULTILAYER_SMALL[-1][0], xstart=0.3e-3, ystart=0.18e-3, speed=1.0,
                  power=250.0, radius=50e-6, powder=POWDER_MULTILAYER_SMALL, toolpath=TOOL_MULTILAYER_SMALL)),
}


def write_case(name, outdir, **overrides):
    """Write <outdir>/<name>.k and <outdir>/<name>_mesh.k; return the deck path."""
    spec = CASES[name]
    os.makedirs(outdir, exist_ok=True)
    mesh_path = os.path.abspath(os.path.join(outdir, name + "_mesh.k"))
    write_mesh(mesh_path, *spec["mesh"])
    deck_path = os.path.join(outdir, name + ".k")
    kw = dict(spec["deck"])
    kw.update(overrides)
    write_deck(deck_path, mesh_path, keyword=name, **kw)
    return deck_path


def sweep_instance(s):
    """Instance s of the 64-instance parameter sweep: drawn with numpy's PCG64
    seeded 20261203 + s, P ~ U(150, 400) W, v ~ U(0.5, 1.5) m/s,
    r ~ U(40, 80) um, nx ~ U{120..320}, ny ~ U{48..128}, nz ~ U{32..64}
    uniform 12.5 um cells, dt = 2e-6 s; the track starts a fifth of the way
    along x on the centre line and runs for min(2 ms, the time to cross
    three fifths of the plate).  Returns a CASES-style dict."""
    import numpy as np
    rng = np.random.default_rng(20261203 + s)
    power = float(rng.uniform(150.0, 400.0))
    speed = float(rng.uniform(0.5, 1.5))
    radius = float(rng.uniform(40e-6, 80e-6))
    nx, ny, nz = (int(rng.integers(lo, hi + 1)) for lo, hi in ((120, 320), (48, 128), (32, 64)))
    h = 12.5e-6
    lx, ly, lz = nx * h, ny * h, nz * h
    xstart = 0.2 * lx
    lasertime = min(2e-3, 0.6 * lx / speed)
    return dict(mesh=(([lx], [nx], [1.0]), ([ly], [ny], [1.0]), ([lz], [nz], [1.0])),
                deck=dict(finaltime=lasertime, maxit=100, dt=2e-6, lasertime=lasertime, xstart=xstart,
                          ystart=0.5 * ly, speed=speed, power=power, radius=radius))


def write_sweep_instance(s, outdir):
    """Write instance s of the sweep as <outdir>/sweep<s>.k; return the path."""
    name = "sweep%02d" % s
    CASES[name] = sweep_instance(s)
    return write_case(name, outdir)
