"""The five workloads of the benchmark, as data: molecules, forcefield
options and run lengths, shared by bench.py and the configuration tests so
that both run exactly the same inputs.

  single   one 20-residue peptide, 12 A cutoff + 2 A skin (the golden
           fixtures c1_seed1..3 hold the reference's own runs of it)
  chain    one 500-residue chain (~8,000 atoms), backbone phi/psi from the
           alpha and beta basins, 12 A cutoff; clash relief, then
           velocity-Verlet MD at 300 K
  batch    4,096 conformers of 10-150 residues, 10 A cutoff, one batched
           energy + force evaluation per step
  serving  minimisation requests of 20-400 residues (log-uniform), 10 A
           cutoff, steepest descent to force RMS < 0.1 kcal/mol/A or 1,000
           steps
  allpairs one 2,000-residue chain (~32,000 atoms), Coulomb and Still GB over
           every pair, LJ at 12 A, velocity-Verlet MD

Molecules come from pdgpu.synth (seeded; see its docstring for what it
builds), or, after use_parameter_file(path) with a PD parameter file, from
PD's own builder (pdgpu.prep: NewProtein of the same sequences, the
chain's phi/psi drawn from the same basins), so that the workloads run on
the reference's molecules wherever its parameter file is at hand."""
import numpy as np

from . import PDGPU_ALL, default_options, synth  # noqa: F401

_FF_PATH = None
_FF_CACHE = {}


def use_parameter_file(path):
    """Build every later workload molecule with PD's NewProtein from this
    parameter file (None: back to pdgpu.synth)."""
    global _FF_PATH
    _FF_PATH = None if path is None else str(path)


def _ff(path):
    from . import prep
    if path not in _FF_CACHE:
        _FF_CACHE[path] = prep.read_ff(path)
    return _FF_CACHE[path]


def basin_angles(seed, n):
    """phi/psi/omega (rad) per residue: 60 % alpha basin (-63, -43), 40 %
    beta (-120, 130), 15 degree spread, trans peptide bonds."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 3))
    for i in range(n):
        if rng.random() < 0.6:
            out[i, :2] = -63.0 + rng.normal(0, 15.0), -43.0 + rng.normal(0, 15.0)
        else:
            out[i, :2] = -120.0 + rng.normal(0, 15.0), 130.0 + rng.normal(0, 15.0)
    out[:, :2] *= np.pi / 180.0
    out[:, 2] = np.pi
    return out


def _pd_job(job):
    from . import prep
    path, seed, residues, mode = job
    seq = synth.random_sequence(seed, residues)
    phipsi = basin_angles(seed, residues) if mode == "basins" else None
    return _ff(path).peptide(prep.pd_grammar(seq), phipsi)


def _many(seeds, residues, mode="extended", workers=0, bonded=False, start="fork"):
    if _FF_PATH is None:
        return synth.many(seeds, residues, mode=mode, workers=workers, bonded=bonded, start=start)
    import multiprocessing as mp
    import os
    jobs = [(_FF_PATH, s, r, mode) for s, r in zip(seeds, residues)]
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(jobs) < 8:
        return [_pd_job(j) for j in jobs]
    with mp.get_context(start).Pool(workers) as pool:
        return pool.map(_pd_job, jobs, chunksize=max(1, len(jobs) // (4 * workers)))

KB = 1.380649e-23  # J/K

# forcefield knobs common to every workload: PD's amber99 1-4 scalings,
# Still GB with eps_in 1, eps_out 78.5, InnerCutoff = Cutoff - 3 A,
# VdwCutoff = Cutoff, VdwInnerCutoff = InnerCutoff
_COMMON = dict(vdw14_scaling=0.5, elec14_scaling=0.83333, dielectric_solvent=78.5, nlist_padding=2.0)


def options(name, precision, **extra):
    """pdgpu options of a workload (precision: pdgpu.FP64 or pdgpu.MIXED)."""
    if name == "single":
        kw = dict(cutoff=12.0, inner_cutoff=9.0, vdw_cutoff=12.0, vdw_inner_cutoff=9.0)
    elif name == "chain":
        # ~1,500 tiles for the ~8,000-atom chain: two wavefronts per tile
        # (measured 0.118 -> 0.087 ms per mixed MD step; 4 and 8 were slower)
        kw = dict(cutoff=12.0, inner_cutoff=9.0, vdw_cutoff=12.0, vdw_inner_cutoff=9.0, tile_split=2)
    elif name in ("batch", "serving"):
        kw = dict(cutoff=10.0, inner_cutoff=7.0, vdw_cutoff=10.0, vdw_inner_cutoff=7.0)
    elif name == "allpairs":
        # Cutoff (Coulomb + the GB pair term, which PD's fast path evaluates
        # together) and the Still cutoff beyond any distance in the molecule,
        # so neither is switched or truncated; LJ at 12 A, switched from 9 A
        kw = dict(cutoff=1.0e4, inner_cutoff=1.0e4, vdw_cutoff=12.0, vdw_inner_cutoff=9.0,
                  gbsa_still_cutoff=1.0e4)
    else:
        raise KeyError(name)
    kw = dict(_COMMON, **kw)
    kw.update(extra)
    return default_options(precision=precision, **kw)


def oracle_params(name, **extra):
    """The same knobs under the oracle's field names (tests only)."""
    o = options(name, 0)
    d = dict(cutoff=o.cutoff, inner_cutoff=o.inner_cutoff, vdw_cutoff=o.vdw_cutoff,
             vdw_inner_cutoff=o.vdw_inner_cutoff, vdw14_scaling=o.vdw14_scaling, elec14_scaling=o.elec14_scaling,
             dielectric_solvent=o.dielectric_solvent, still_cutoff=o.gbsa_still_cutoff, padding=o.nlist_padding)
    d.update(extra)
    return d


def batch_lengths(n=4096):
    return [int(L) for L in np.random.default_rng(11).integers(10, 151, size=n)]


def batch_molecules(n=4096, workers=0, start="fork"):
    """The batch: n extended conformers, lengths uniform in 10-150 residues,
    conformer k built from seed 1000 + k."""
    return _many([1000 + k for k in range(n)], batch_lengths(n), workers=workers, start=start)


def chain_molecule():
    if _FF_PATH is not None:
        return _pd_job((_FF_PATH, 2, 500, "basins"))
    return synth.peptide(seed=2, residues=500, mode="basins", bonded=True)


def serving_sizes(n=200):
    return [int(x) for x in np.exp(np.random.default_rng(17).uniform(np.log(20), np.log(400), size=n))]


def serving_molecules(n=200, workers=0, start="fork"):
    """n requests, sizes log-uniform in 20-400 residues, request k built
    from seed 5000 + k, with FF_Bonded terms."""
    return _many([5000 + k for k in range(n)], serving_sizes(n), bonded=True, workers=workers, start=start)


def allpairs_molecule(residues=2000):
    if _FF_PATH is not None:
        return _pd_job((_FF_PATH, 2000, residues, "extended"))
    return synth.peptide(seed=2000, residues=residues, bonded=True)


def maxwell(mass, temperature, seed):
    """Maxwell-Boltzmann velocities (m/s) without net momentum."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(len(mass), 3)) * np.sqrt(KB * temperature / mass)[:, None]
    v -= (mass[:, None] * v).sum(0) / mass.sum()
    return v


def temperature(mass, vel):
    """Instantaneous temperature (K) of velocities vel (m/s), masses in kg."""
    return float((mass[:, None] * vel ** 2).sum() / (3.0 * len(mass) * KB))


def equilibrate(ctx, iid, mass, temperature_k, rounds=10, every=100, dt=1e-15):
    """Velocity-rescaling equilibration before an NVE run: `rounds` times,
    `every` velocity-Verlet steps then velocities scaled to temperature_k.
    From a minimised structure, half of the starting kinetic energy flows
    into the potential within ~100 fs; the rescaling puts it back, so the
    run that follows holds temperature_k instead of ~temperature_k / 2."""
    for _ in range(rounds):
        ctx.md_vv(iid, every, dt, 10)
        v = ctx.get_velocities(iid)
        ctx.set_velocities(iid, v * np.sqrt(temperature_k / temperature(mass, v)))


def relax(ctx_factory, mol, steps=4000):
    """Clash relief of a synthetic chain: `steps` of PD's steepest descent
    (StepSize 2e5, list rebuilt every step) over FF_Bonded + GB in mixed
    precision; returns the positions. For the 500-residue chain: +5.1e5
    kcal/mol as built, +1.6e4 after 100 steps, -2.2e3 after 4,000; velocity
    Verlet from 300 K Maxwell velocities then holds ~235 K over 10,000
    steps (after 100 steps it heated to ~600 K)."""
    with ctx_factory() as ctx:
        i = ctx.add_system(mol, bonded=True)
        ctx.minimise_sd(i, steps, 2e5, 1)
        return ctx.get_positions(i)
