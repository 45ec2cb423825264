"""Shared builder for the tiny TRANSPORTSE configs behind the golden vectors.

Each ``g*.py`` config module calls ``build(globals(), ...)``, which fills the
module namespace with every attribute the reference's config contract reads
(tse.py:41-167, 1120-1172): grid, flags, fields, callbacks and solver
settings. The physics follows the benchmark cases at a size the pure-Python
reference finishes in seconds: variable-density Darcy flow with a linear EOS
rho = 1000 + 700*C_salt, an optional heat equation with conductivity linear
in T, lognormal permeability, geometric z stretch and slope angle beta.

When the environment variable TSE_GOLDEN_DUMP names a directory, the
``dynamic_bc`` callback (called once per time step, tse.py:1449-1453) saves
the solver state of the running ``__main__`` module there; it changes
nothing otherwise, unless ``dyn_inplace`` asks it to scale phi and nue in
place.
"""
import os
import sys

import numpy


def build(ns, nx=7, ny=6, nz=5, seed=11, heat=False, species=1, sptrans=(0,),
          sia=False, sia_maxit=1, sia_conv=1e-4, c_imblflo=True,
          t_imblflo=False, omega_c=1, omega_t=1, sor_p=1.8, sor_c=1.0,
          sor_t=1.0, conv_p=1e-8, conv_c=1e-8, conv_t=1e-8, nsteps=4,
          dt0=86400.0, dt_max=86400.0, beta=5.0, stretch=1.05,
          maxit_p=0, maxit_c=0, maxit_t=0, disp=(1.0, 0.1), wp_cell=True,
          plot_steps=0, c_f=4.5e-10, salt_patch=True, t_end=None,
          smolar_c=False, smolar_t=False, nlte=False, bp_faces="x",
          dyn_inplace=False, plot_times=None):
    rng = numpy.random.default_rng(seed)
    S = species
    ns.update(dict(
        parallel_processes=1, parallel_chemistry=1, numba_fastmath=False,
        numba_tl='omp', nx=nx, ny=ny, nz=nz,
        dx=numpy.ones(nx) * 10.0, dy=numpy.ones(ny) * 10.0,
        dz=2.0 * stretch ** numpy.arange(nz)[::-1],
        g=9.80665, beta=beta, t=dt0 * nsteps if t_end is None else t_end,
        FLUIDFLOW=True, SPECTRANS=True, HEATTRANS=heat, FLUIDPROP=True,
        CHEMREACT=False, DYNAMICBC=True, MASSBALAN=True, DTCHEM=1,
        C_IMBLFLO=c_imblflo, T_IMBLFLO=t_imblflo, SIA=sia,
        SIA_maxit=sia_maxit, SIA_conv=sia_conv, SMOLAR_C=smolar_c,
        SMOLAR_T=smolar_t, NLTE_SOLV=nlte, SPECIES=S,
        SPTRANS=numpy.array(sptrans),
        dt=dt0, dt_max=dt_max, OMEGA_C=omega_c, OMEGA_T=omega_t,
        SOR_P=sor_p, SOR_C=sor_c, SOR_T=sor_t, CONV_P=conv_p,
        CONV_C=conv_c, CONV_T=conv_t, PRINTOUT=1,
        PLOTSAVE=plot_times is not None,
        PLOTTIME=numpy.array([0.0] if plot_times is None else plot_times,
                             dtype=float),
        MAXIT_P=maxit_p, MAXIT_C=maxit_c,
        MAXIT_T=maxit_t))
    ns['lx'] = float(ns['dx'].sum())
    ns['ly'] = float(ns['dy'].sum())
    ns['lz'] = float(ns['dz'].sum())
    shp = (nx, ny, nz)
    K = 1e-12 * numpy.exp(1.0 * rng.standard_normal(shp))
    ns['phi'] = numpy.full(shp, 0.2)
    ns['Kx'] = K.copy()
    ns['Ky'] = K.copy()
    ns['Kz'] = K.copy()
    ns['c_r'] = numpy.full(shp, 1e-10)
    ns['cp_r'] = numpy.full(shp, 1000.0)
    ns['l_r'] = numpy.full(shp, 2.0)
    ns['rho_r'] = numpy.full(shp, 2500.0)
    ns['lf_l'] = numpy.full(shp, disp[0])
    ns['lf_t'] = numpy.full(shp, disp[1])
    ns['lt_l'] = numpy.full(shp, disp[0])
    ns['lt_t'] = numpy.full(shp, disp[1])
    ns['L'] = numpy.zeros(shp)
    ns['R'] = numpy.ones(shp)
    ns['mass'] = numpy.zeros(3)
    Df = numpy.zeros(shp + (S,))
    for s in range(S):
        Df[..., s] = 1e-9 * (1.0 + 0.5 * s)
    ns['D_fx'] = Df.copy()
    ns['D_fy'] = Df.copy()
    ns['D_fz'] = Df.copy()
    ns["c_f"] = numpy.full(shp, c_f)
    ns['cp_f'] = numpy.full(shp, 4150.0)
    ns['l_f'] = numpy.full(shp, 2.5)
    ns['rho_f'] = numpy.full(shp, 1000.0)
    ns['nue'] = numpy.full(shp, 1e-3)
    # salt: C = 1 in the top two layers over the central x third
    C = numpy.zeros(shp + (S,))
    if salt_patch:
        C[nx // 3: nx - nx // 3, :, nz - 2:, 0] = 1.0
    if S > 1:
        C[:, :, :, 1] = rng.uniform(0.0, 0.1, shp)
    ns['C'] = C
    # NLTE: layer 0 fluid, layer 1 rock (tse.py:1343-1354)
    ns['T'] = numpy.full(shp + (2,), 10.0) if nlte else numpy.full(shp, 10.0)
    ns['BP'] = numpy.zeros(shp)
    ns['BC'] = numpy.zeros(shp)
    ns['BT'] = numpy.zeros(shp)
    if bp_faces == "x":
        ns['BP'][0, :, :] = 1
        ns['BP'][-1, :, :] = 1
    else:  # columns: Dirichlet top and bottom
        ns['BP'][:, :, 0] = 1
        ns['BP'][:, :, -1] = 1
    if heat:
        ns['BT'][:, :, 0] = 1
        ns['T'][:, :, 0] = 60.0
        if nlte:  # rock layer starts warmer than the fluid in the middle
            ns['T'][nx // 2, :, nz // 2, 1] = 30.0
    ns['WP'] = numpy.zeros(shp)
    ns['WT'] = numpy.zeros(shp)
    ns['WC'] = numpy.zeros(shp + (S,))
    if wp_cell:
        kw = min(1, nz - 1)
        ns['WP'][nx // 2, ny // 2, kw] = 2e-6
        ns['WC'][nx // 2, ny // 2, kw, 0] = 1e-7
        if heat:
            ns['WT'][nx // 2, ny // 2, kw] = 5.0
    ns['qx'] = numpy.zeros(shp)
    ns['qy'] = numpy.zeros(shp)
    ns['qz'] = numpy.zeros(shp)
    # hydrostatic pressure in the tilted frame (k points up, top at nz-1)
    dz = ns['dz']
    b = numpy.radians(beta)
    P = numpy.zeros(shp)
    col = numpy.zeros(nz)
    col[nz - 1] = 1e5 + 1000.0 * 9.80665 * numpy.cos(b) * 0.5 * dz[nz - 1]
    for k in range(nz - 2, -1, -1):
        col[k] = col[k + 1] + 1000.0 * 9.80665 * numpy.cos(b) * 0.5 * (dz[k] + dz[k + 1])
    xs = numpy.cumsum(ns['dx']) - 0.5 * ns['dx']
    for i in range(nx):
        P[i, :, :] = col[None, :] - 1000.0 * 9.80665 * numpy.sin(b) * xs[i]
    ns['P'] = P

    def fluid_properties(C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi,
                         Kx, Ky, Kz, D_fx, D_fy, D_fz, rho_r):
        rho_f = 1000.0 + 700.0 * C[0]
        l_f = 2.5 * (1.0 + 0.002 * (T - 10.0))
        return (C, T, P, nue, rho_f, c_f, cp_f, l_f, cp_r, l_r, phi, Kx, Ky,
                Kz, D_fx, D_fy, D_fz, rho_r)

    def chemical_reactions(C, P, T, rho_f, phi, Kx, Ky, Kz, dt_chem, i, j, k):
        return C, P, T, rho_f, phi, Kx, Ky, Kz

    vol = (ns['dx'][:, None, None] * ns['dy'][None, :, None]
           * ns['dz'][None, None, :])

    def mass_balance(C, qx, qy, qz, WP, WC, WT, D_fx, D_fy, D_fz, dt, mass):
        mass = mass.copy()
        mass[0] = numpy.sum(C[:, :, :, 0] * ns['phi'] * vol)
        mass[1] += dt * numpy.sum(numpy.abs(qx[max(nx - 2, 0), :, :]) * ns['dy'][:, None]
                                  * ns['dz'][None, :])
        mass[2] += dt * numpy.sum(WC[:, :, :, 0] * vol)
        return mass

    def dynamic_bc(t, bc_inc, P, T, C, BP, BT, BC, WP, WT, WC, Kx, Ky, Kz,
                   rho_f, nue, c_f, cp_f, qx, phi, l_f, l_r):
        out = os.environ.get('TSE_GOLDEN_DUMP')
        if out:
            m = sys.modules['__main__']
            rec = dict(t_end=t, dt=m.dt, P=P, T=T, C=C, qx=qx,
                       qy=m.qy, qz=m.qz, e=m.e, rho_f=rho_f,
                       l_f=l_f, numitP=m.numitP, numitC=m.numitC,
                       numitT=m.numitT)
            for name in ('dt_P', 'dt_C', 'dt_T'):
                rec[name] = getattr(m, name, numpy.nan)
            numpy.savez(os.path.join(out, 'step%05d.npz' % m.n_tsteps), **rec)
        if dyn_inplace:
            # arrays the callback receives but does not return, changed in
            # place: a caller that caches device copies must see the change
            phi *= 1.02
            nue *= 0.99
        return bc_inc, P, T, C, BP, BT, BC, WP, WT, WC, Kx, Ky, Kz

    ns['fluid_properties'] = fluid_properties
    ns['chemical_reactions'] = chemical_reactions
    ns['mass_balance'] = mass_balance
    ns['dynamic_bc'] = dynamic_bc
