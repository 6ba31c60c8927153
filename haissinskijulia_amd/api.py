"""Host-side mirror of HaissinskiJulia's tracking API over the MI355X library.

Names, argument order and return values follow the reference
(src/StochasticHaissinski.jl:46-56 exports):

* ``SimulationParameters`` -- src/data_structures.jl:81-110 (field order kept;
  ``α_c``, ``ϕs``, ``update_η`` are also accepted by those names).
* ``Particles`` with ``.coordinates.z`` / ``.coordinates.ΔE`` -- the SoA
  ``StructArray{Particle{Float64}}`` of src/data_structures.jl:26-42.
* ``create_simulation_buffers(n, nbins, use_mpi, comm_size)`` --
  src/utils.jl:147-254; only ``nbins = length(buffers.λ)`` and the QE draw
  buffer are used.
* ``longitudinal_evolve(particles, params, buffers, comm, use_mpi)`` --
  ``longitudinal_evolve!`` of src/evolution.jl:206-212; mutates the particles
  and returns ``(σ_E, σ_z, E0)``.  A ``MI355X`` token in the ``comm`` slot
  selects the GPU (one GPU holds the whole bunch, so the MPI collectives of
  the reference become on-device reductions).  ``use_mpi = False`` selects
  the reference's serial branch: two-pass ``compute_std`` spreads, no E0
  update of an empty bunch, and the reference's error with the wake on.
* stage functions ``rf_kick``, ``quantum_excitation``,
  ``synchrotron_radiation``, ``apply_wakefield_inplace`` and the helpers
  ``z_to_ϕ``, ``ϕ_to_z``, ``calc_rf_factor``, ``next_power_of_two``.

Errors carry the reference's message text (evolution.jl:264,278) and the
library status code.
"""
from __future__ import annotations

import atexit
import ctypes as C
import math
import weakref
from dataclasses import dataclass, field, fields
from typing import Optional, Sequence

import numpy as np

from . import _lib
from . import tune as _tune
from ._lib import HaissinskiError, HmOptions, HmParams, HmResult, check, dptr, iptr

SPEED_LIGHT = 299792458.0          # src/constants.jl:5
ELECTRON_CHARGE = 1.602176634e-19  # src/constants.jl:6
MASS_ELECTRON = 0.51099895069e6    # src/constants.jl:7

_ALIASES = {"α_c": "alpha_c", "ϕs": "phi_s", "φs": "phi_s", "update_η": "update_eta"}


# ---------------------------------------------------------------------------
# containers
# ---------------------------------------------------------------------------

@dataclass
class SimulationParameters:
    """src/data_structures.jl:81-97 (positional order of the reference)."""
    E0: float
    mass: float
    voltage: float
    harmonic: int
    radius: float
    pipe_radius: float
    alpha_c: float
    phi_s: float
    freq_rf: float
    n_turns: int
    use_wakefield: bool
    update_eta: bool
    update_E0: bool
    SR_damping: bool
    use_excitation: bool

    @classmethod
    def from_dict(cls, d: dict) -> "SimulationParameters":
        kw = {_ALIASES.get(k, k): v for k, v in d.items()}
        return cls(**{f.name: kw[f.name] for f in fields(cls)})

    def replace(self, **kw) -> "SimulationParameters":
        d = {f.name: getattr(self, f.name) for f in fields(self)}
        d.update({_ALIASES.get(k, k): v for k, v in kw.items()})
        return SimulationParameters(**d)

    def as_dict(self) -> dict:
        return {f.name: getattr(self, f.name) for f in fields(self)}

    def _c(self) -> HmParams:
        return HmParams(self.E0, self.mass, self.voltage, int(self.harmonic), self.radius,
                        self.pipe_radius, self.alpha_c, self.phi_s, self.freq_rf,
                        int(self.n_turns), int(bool(self.use_wakefield)), int(bool(self.update_eta)),
                        int(bool(self.update_E0)), int(bool(self.SR_damping)),
                        int(bool(self.use_excitation)))


for _uni, _ascii in (("α_c", "alpha_c"), ("φs", "phi_s"), ("update_η", "update_eta")):
    setattr(SimulationParameters, _uni, property(lambda self, a=_ascii: getattr(self, a)))


class Coordinates:
    """Field storage of StructArray{Coordinate}: contiguous FP64 z and ΔE."""

    def __init__(self, z: np.ndarray, dE: np.ndarray):
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        self.ΔE = np.ascontiguousarray(dE, dtype=np.float64)
        if self.z.shape != self.ΔE.shape or self.z.ndim != 1:
            raise ValueError("z and ΔE must be 1-d arrays of one length")

    @property
    def dE(self) -> np.ndarray:
        return self.ΔE


class Particles:
    """StructArray{Particle{Float64}} stand-in (src/data_structures.jl:39-42)."""

    def __init__(self, z, dE):
        self.coordinates = Coordinates(z, dE)

    def __len__(self) -> int:
        return self.coordinates.z.size

    def copy(self) -> "Particles":
        return Particles(self.coordinates.z.copy(), self.coordinates.ΔE.copy())


@dataclass
class SimulationBuffers:
    """The parts of SimulationBuffers (src/data_structures.jl:141-192) the GPU
    path reads: the bin count (length of λ) and the QE draw buffer."""
    λ: np.ndarray
    random_buffer: np.ndarray
    use_mpi: bool = True
    comm_size: int = 1

    @property
    def nbins(self) -> int:
        return int(self.λ.size)


def create_simulation_buffers(n_particles: int, nbins: int, use_mpi: bool = True,
                              comm_size: int = 1) -> SimulationBuffers:
    """src/utils.jl:147-160 (the same argument checks)."""
    if nbins <= 0:
        raise HaissinskiError(_lib.HM_ERR_NBINS, f"Number of bins (nbins) must be positive. Got {nbins}.")
    if n_particles < 0:
        raise HaissinskiError(_lib.HM_ERR_ARG,
                              f"Number of particles must be non-negative. Got {n_particles}.")
    return SimulationBuffers(np.zeros(nbins), np.zeros(n_particles), use_mpi, comm_size)


@dataclass
class TrackingOptions:
    """Knobs outside SimulationParameters (all defaults = the reference)."""
    seed: int = 0                        # Philox key of the QE draws
    turn_offset: int = 0                 # absolute index of the first turn
    normals: Optional[np.ndarray] = None  # injected QE normals [n_turns, n]
    wake_table: Optional[np.ndarray] = None   # W by bin lag [nbins]
    resonators: Sequence[tuple] = ()     # (R_s [Ohm], f_r [Hz], Q) broadband resonators
    delta_sigma: float = 0.0             # 0 -> 0.15 (wakefield.jl:204); <0 -> no smoothing kernel
    bin_half_width_sigmas: float = 0.0   # 0 -> 7.5 (evolution.jl:270)
    charge: float = 0.0                  # >0 -> wake current = charge [C]
    timing: bool = False
    fft_block: int = 0
    block_size: int = 0
    monitor: bool = False                # record per-turn centroid, sigma_z, E0, binned count
    serial: bool = False                 # the use_mpi = false branch (longitudinal_evolve sets it)
    fom_stats: bool = False              # final two-pass means and cor(z, ΔE) (figures of merit)
    tile_pairs: int = 0                  # pairs per thread per tile of the turn pass (0: tuned/rule)
    run_pairs: int = 0                   # cost of a run of tiles when balancing the pass (0: tuned/rule)
    bin_whole: int = 0                   # bin phase of a small bunch in one workgroup: 0 rule, -1 never

    def _c(self, keep: list, n: int, n_turns: int, nbins: int) -> HmOptions:
        o = HmOptions()
        o.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        o.turn_offset = int(self.turn_offset)
        if self.normals is not None:
            nrm = np.ascontiguousarray(self.normals, dtype=np.float64).reshape(-1)
            if nrm.size < n * n_turns:
                raise ValueError(f"normals holds {nrm.size} draws, needs n_turns*n = {n * n_turns}")
            keep.append(nrm)
            o.normals = dptr(nrm)
        if self.wake_table is not None:
            wt = np.ascontiguousarray(self.wake_table, dtype=np.float64)
            if wt.size < nbins:
                raise ValueError("wake_table shorter than nbins")
            keep.append(wt)
            o.wake_table = dptr(wt)
        if self.resonators:
            rs = np.ascontiguousarray(np.asarray(self.resonators, dtype=np.float64).reshape(-1))
            keep.append(rs)
            o.n_resonators = rs.size // 3
            o.resonators = dptr(rs)
        o.delta_sigma = self.delta_sigma
        o.bin_half_width_sigmas = self.bin_half_width_sigmas
        o.charge = self.charge
        o.timing = int(bool(self.timing))
        # 0: the tuned block for this bin count (tuning.json), else the library's rule
        o.fft_block = int(self.fft_block) or _tune.fft_block_for(nbins)
        o.block_size = int(self.block_size)
        if self.monitor and n_turns > 0:
            mon = np.zeros((n_turns, 4))
            keep.append(mon)
            o.monitor = dptr(mon)
            o._monitor = mon
        o.serial = int(bool(self.serial))
        o.fom_stats = int(bool(self.fom_stats))
        # 0: the tuned values for this bunch size and bin count (tuning.json),
        # else the library's rules
        o.tile_pairs = int(self.tile_pairs) or _tune.tile_pairs_for(n)
        o.run_pairs = int(self.run_pairs) or _tune.run_pairs_for(nbins)
        o.bin_whole = int(self.bin_whole)
        return o


@dataclass
class EvolveResult:
    sigma_E: float
    sigma_z: float
    E0: float
    turns_done: int
    drift_skipped: int
    ms_total: float
    ms_pass_avg: float
    ms_conv_avg: float
    pass_launches: int
    sigma_E0: float
    sigma_z0: float
    bin_half_width: float
    mean_z: float = 0.0
    mean_dE: float = 0.0
    corr_z_dE: float = 0.0    # Statistics.cor of the final bunch (two-pass, on the GPU;
                              # NaN unless TrackingOptions.fom_stats)
    # per turn (TrackingOptions.monitor): z centroid and sigma_z after the drift
    # (wake on), E0 after the turn, binned count
    monitor: Optional[np.ndarray] = None

    @classmethod
    def _from(cls, r: HmResult) -> "EvolveResult":
        return cls(r.sigma_E, r.sigma_z, r.E0, r.turns_done, r.drift_skipped, r.ms_total,
                   r.ms_pass_avg, r.ms_conv_avg, r.pass_launches, r.sigma_E0, r.sigma_z0,
                   r.bin_half_width, r.mean_z, r.mean_dE, r.corr_z_dE)

    def __iter__(self):
        """Unpacks like the reference's return value: (σ_E, σ_z, E0)."""
        return iter((self.sigma_E, self.sigma_z, self.E0))


# ---------------------------------------------------------------------------
# device token and device-resident bunches
# ---------------------------------------------------------------------------

class MI355X:
    """Device token for the ``comm`` slot: one context (GPU + stream).

    Bunches and servers on the context are released before it (``close``
    releases them first), and every context still open when the interpreter
    exits is released by an ``atexit`` hook, while the HIP runtime is whole."""

    _default: "MI355X | None" = None

    def __init__(self, device: int = 0):
        self.device = device
        self._ctx = None
        self._dependents = weakref.WeakSet()   # bunches and servers on this context

    @property
    def ctx(self):
        if self._ctx is None:
            L = _lib.load()
            h = C.c_void_p()
            check(L.hm_create(self.device, C.byref(h)))
            self._ctx = h
            _open_devices.add(self)
        return self._ctx

    def info(self) -> dict:
        arch = C.create_string_buffer(64)
        cus = C.c_int32()
        lds = C.c_int32()
        check(_lib.load().hm_device_info(self.ctx, arch, 64, C.byref(cus), C.byref(lds)))
        return {"arch": arch.value.decode(), "cus": cus.value, "lds_bytes": lds.value}

    def close(self) -> None:
        if self._ctx is not None:
            for d in list(self._dependents):   # they live on the context
                d.close()
            _lib.load().hm_destroy(self._ctx)
            self._ctx = None

    @classmethod
    def default(cls) -> "MI355X":
        if cls._default is None:
            cls._default = MI355X(0)
        return cls._default

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# contexts with an open hm_ctx; released at interpreter exit, each after its
# bunches and servers (during module teardown a __del__ may find this module
# or the runtime already gone)
_open_devices: "weakref.WeakSet[MI355X]" = weakref.WeakSet()


@atexit.register
def _release_open_devices() -> None:
    for dev in list(_open_devices):
        try:
            dev.close()
        except Exception:
            pass


def _dev(comm) -> MI355X:
    if isinstance(comm, MI355X):
        return comm
    if comm is None:
        return MI355X.default()
    raise TypeError(f"comm must be an MI355X device token (got {type(comm).__name__})")


class DeviceBunch:
    """Particles resident in HBM (hm_bunch)."""

    def __init__(self, n: int, device: Optional[MI355X] = None):
        self.device = device or MI355X.default()
        self.n = int(n)
        h = C.c_void_p()
        check(_lib.load().hm_bunch_create(self.device.ctx, self.n, C.byref(h)))
        self._h = h
        self.device._dependents.add(self)

    @classmethod
    def from_particles(cls, particles: Particles, device: Optional[MI355X] = None) -> "DeviceBunch":
        b = cls(len(particles), device)
        b.upload(particles.coordinates.z, particles.coordinates.ΔE)
        return b

    def upload(self, z: np.ndarray, dE: np.ndarray) -> None:
        z = np.ascontiguousarray(z, dtype=np.float64)
        dE = np.ascontiguousarray(dE, dtype=np.float64)
        if z.size != self.n or dE.size != self.n:
            raise ValueError("size mismatch")
        check(_lib.load().hm_bunch_upload(self._h, dptr(z), dptr(dE)))

    def download(self):
        z = np.empty(self.n)
        dE = np.empty(self.n)
        check(_lib.load().hm_bunch_download(self._h, dptr(z), dptr(dE)))
        return z, dE

    def generate(self, seed: int, sigma_z: float, sigma_E: float, mu_z: float = 0.0,
                 mu_E: float = 0.0) -> None:
        """Independent Gaussian z and dE from the Philox stream of `seed`."""
        check(_lib.load().hm_bunch_generate(self._h, seed, mu_z, sigma_z, mu_E, sigma_E))

    def generate_particles(self, seed: int, mu_z: float, mu_E: float, sigma_z: float,
                           sigma_E: float) -> tuple[float, float]:
        """generate_particles (src/evolution.jl:42-133) into this bunch;
        returns (σ_E, σ_z) as the reference does."""
        sE, sz = C.c_double(), C.c_double()
        check(_lib.load().hm_bunch_generate_particles(self._h, seed, mu_z, mu_E, sigma_z, sigma_E,
                                                      C.byref(sE), C.byref(sz)))
        return sE.value, sz.value

    def evolve(self, params: SimulationParameters, nbins: int,
               options: Optional[TrackingOptions] = None) -> EvolveResult:
        options = options or TrackingOptions()
        keep: list = []
        o = options._c(keep, self.n, params.n_turns, nbins)
        r = HmResult()
        check(_lib.load().hm_bunch_evolve(self._h, nbins, C.byref(params._c()), C.byref(o), C.byref(r)))
        res = EvolveResult._from(r)
        res.monitor = getattr(o, "_monitor", None)
        return res

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().hm_bunch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_evolve(bunches: Sequence[DeviceBunch], nbins: Sequence[int],
                 params: Sequence[SimulationParameters],
                 options: Optional[Sequence[TrackingOptions]] = None) -> list[EvolveResult]:
    """Advance several resident bunches together (one segmented launch per turn)."""
    nb = len(bunches)
    if not (len(nbins) == len(params) == nb) or (options is not None and len(options) != nb):
        raise ValueError("bunches, nbins, params and options must have one length")
    dev = bunches[0].device
    keep: list = []
    hs = (C.c_void_p * nb)(*[b._h for b in bunches])
    nbs = np.ascontiguousarray(nbins, dtype=np.int64)
    ps = (HmParams * nb)(*[p._c() for p in params])
    opts = options or [TrackingOptions() for _ in range(nb)]
    cs = [o._c(keep, b.n, p.n_turns, k) for o, b, p, k in zip(opts, bunches, params, nbins)]
    os_ = (HmOptions * nb)(*cs)
    rs = (HmResult * nb)()
    check(_lib.load().hm_batch_evolve(dev.ctx, nb, hs, iptr(nbs), ps, os_, rs))
    out = [EvolveResult._from(r) for r in rs]
    for res, c in zip(out, cs):
        res.monitor = getattr(c, "_monitor", None)
    return out


@dataclass
class Job:
    """One tracking job for ``Server``: a Gaussian bunch generated on the GPU
    from ``particle_seed`` (``DeviceBunch.generate``), then ``params.n_turns``
    turns."""
    params: SimulationParameters
    n: int
    nbins: int
    particle_seed: int
    sigma_z: float
    sigma_E: float
    options: TrackingOptions = field(default_factory=TrackingOptions)
    fail_after_turns: int = 0   # >0: stop the job early with HM_ERR_CANCELLED (failure hook)
    fail_setup: bool = False    # its set-up inside the batch fails (admission error-path hook)


@dataclass
class JobResult:
    job_id: int
    status: int            # 0, or HM_ERR_CANCELLED for cancelled / failed jobs
    result: EvolveResult
    t_submit: float        # server clock [s]
    t_admit: float
    t_done: float
    centroid_z: float
    centroid_E: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class Server:
    """Continuous batching of tracking jobs on one GPU (hm_server_*).

    ``submit`` queues a job; each ``tick(turns)`` admits queued jobs while the
    HBM budget and slot count allow, advances every resident job by up to
    ``turns`` turns in shared launches and retires finished ones; ``poll``
    returns finished results in completion order.  A job's result does not
    depend on what else is resident."""

    def __init__(self, budget_bytes: int, max_slots: int, device: Optional[MI355X] = None):
        self.device = device or MI355X.default()
        h = C.c_void_p()
        check(_lib.load().hm_server_create(self.device.ctx, int(budget_bytes), int(max_slots),
                                           C.byref(h)))
        self._h = h
        self.device._dependents.add(self)

    def submit(self, job: Job) -> int:
        keep: list = []
        j = _lib.HmJob()
        j.params = job.params._c()
        j.opts = job.options._c(keep, job.n, job.params.n_turns, job.nbins)
        j.n = int(job.n)
        j.nbins = int(job.nbins)
        j.particle_seed = int(job.particle_seed) & 0xFFFFFFFFFFFFFFFF
        j.sigma_z = float(job.sigma_z)
        j.sigma_E = float(job.sigma_E)
        j.fail_after_turns = int(job.fail_after_turns)
        j.fail_setup = int(bool(job.fail_setup))
        jid = C.c_int64()
        check(_lib.load().hm_server_submit(self._h, C.byref(j), C.byref(jid)))
        return jid.value

    def submit_voltage_pair(self, job: Job, rel_step: float = 0.005) -> tuple[int, int]:
        """An RF-voltage sensitivity request: two jobs at V (1 -+ rel_step) with
        common random numbers (same particle and QE seeds).  Returns their ids;
        ``fd_sensitivity`` turns the two results into dFoM/dV."""
        import dataclasses
        opts = dataclasses.replace(job.options, fom_stats=True)   # the FoMs need cor(z, ΔE)
        lo = dataclasses.replace(job, options=opts, params=job.params.replace(
            voltage=job.params.voltage * (1.0 - rel_step)))
        hi = dataclasses.replace(job, options=opts, params=job.params.replace(
            voltage=job.params.voltage * (1.0 + rel_step)))
        return self.submit(lo), self.submit(hi)

    def cancel(self, job_id: int) -> None:
        check(_lib.load().hm_server_cancel(self._h, int(job_id)))

    def tick(self, turns: int) -> int:
        """Advance resident jobs; returns the number still queued or running."""
        act = C.c_int32()
        check(_lib.load().hm_server_tick(self._h, int(turns), C.byref(act)))
        return act.value

    def poll(self, cap: int = 256) -> list[JobResult]:
        buf = (_lib.HmJobResult * cap)()
        cnt = C.c_int32()
        check(_lib.load().hm_server_poll(self._h, buf, cap, C.byref(cnt)))
        return [JobResult(r.job_id, r.status, EvolveResult._from(r.result), r.t_submit, r.t_admit,
                          r.t_done, r.centroid_z, r.centroid_E) for r in buf[:cnt.value]]

    def run(self, turns_per_tick: int = 100) -> list[JobResult]:
        """Tick until every submitted job is finished; all results."""
        out: list[JobResult] = []
        while self.tick(turns_per_tick) > 0:
            out += self.poll()
        return out + self.poll()

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().hm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# the reference's entry points
# ---------------------------------------------------------------------------

def longitudinal_evolve(particles: Particles, params: SimulationParameters,
                        buffers: SimulationBuffers, comm=None, use_mpi: bool = True,
                        options: Optional[TrackingOptions] = None):
    """longitudinal_evolve! (src/evolution.jl:206-471) on the GPU.

    Mutates ``particles`` and returns ``(σ_E, σ_z, E0)``.  ``options``
    carries the seed and the extensions; ``options.timing`` adds HIP-event
    timings, available afterwards as ``longitudinal_evolve.last``.
    """
    import dataclasses
    dev = _dev(comm)
    options = dataclasses.replace(options or TrackingOptions(), serial=not use_mpi)
    z = particles.coordinates.z
    dE = particles.coordinates.ΔE
    n = z.size
    nbins = buffers.nbins
    keep: list = []
    o = options._c(keep, n, params.n_turns, nbins)
    r = HmResult()
    check(_lib.load().hm_evolve(dev.ctx, dptr(z), dptr(dE), n, nbins, C.byref(params._c()),
                                C.byref(o), C.byref(r)))
    longitudinal_evolve.last = EvolveResult._from(r)
    longitudinal_evolve.last.monitor = getattr(o, "_monitor", None)
    return r.sigma_E, r.sigma_z, r.E0


longitudinal_evolve.last = None


def rf_kick(voltage, sin_ϕs, rf_factor, ϕs, particles: Particles, buffers=None, comm=None):
    """rf_kick! (src/rf_kick.jl:18-25,84-87)."""
    c = particles.coordinates
    check(_lib.load().hm_rf_kick(_dev(comm).ctx, voltage, sin_ϕs, rf_factor, ϕs, dptr(c.z),
                                 dptr(c.ΔE), c.z.size))


def quantum_excitation(E0, radius, σ_E0, buffers: SimulationBuffers, particles: Particles,
                       comm=None, normals: Optional[np.ndarray] = None, seed: int = 0,
                       turn: int = 0):
    """quantum_excitation! (src/quantum_excitation.jl:24-72).

    The reference fills ``buffers.random_buffer`` with ``randn!``; here the
    draws are ``normals`` if given, else the device Philox stream for
    (seed, particle, turn).  They are left in ``buffers.random_buffer``.
    """
    c = particles.coordinates
    n = c.z.size
    if normals is None:
        normals = draw_normals(seed, turn, 0, n, comm)
    r = np.ascontiguousarray(normals, dtype=np.float64)
    if buffers is not None and buffers.random_buffer.size == n:
        buffers.random_buffer[:] = r
    check(_lib.load().hm_quantum_excitation(_dev(comm).ctx, E0, radius, σ_E0, dptr(r), dptr(c.ΔE), n))


def synchrotron_radiation(E0, radius, particles: Particles, buffers=None, comm=None):
    """synchrotron_radiation! (src/synchrotron_radiation.jl:12-58)."""
    c = particles.coordinates
    check(_lib.load().hm_synchrotron_radiation(_dev(comm).ctx, E0, radius, dptr(c.ΔE), c.z.size))


def drift(particles: Particles, params: SimulationParameters, E0: float, comm=None) -> bool:
    """The phase advance of one turn (evolution.jl:330-347); True if skipped."""
    c = particles.coordinates
    sk = C.c_int32()
    check(_lib.load().hm_drift(_dev(comm).ctx, int(bool(params.update_eta)), E0, params.mass,
                               params.alpha_c, int(params.harmonic), params.phi_s, params.freq_rf,
                               dptr(c.z), dptr(c.ΔE), c.z.size, C.byref(sk)))
    return bool(sk.value)


def histogram(z: np.ndarray, bin_edges: np.ndarray, comm=None) -> np.ndarray:
    """calculate_local_histogram! (src/utils.jl:351-401) on uniform edges."""
    z = np.ascontiguousarray(z, dtype=np.float64)
    nb = len(bin_edges) - 1
    step = (2.0 * bin_edges[-1]) / nb if bin_edges[0] == -bin_edges[-1] else \
        (bin_edges[-1] - bin_edges[0]) / nb
    counts = np.zeros(nb, dtype=np.int64)
    check(_lib.load().hm_histogram(_dev(comm).ctx, dptr(z), z.size, float(bin_edges[0]),
                                   float(bin_edges[-1]), 1.0 / step, nb, iptr(counts)))
    return counts


def wake_potential(counts: np.ndarray, wake_tab: np.ndarray, lam_tab: np.ndarray, current: float,
                   comm=None, fft_block: int = 0) -> np.ndarray:
    """wakefield.jl:186-319: potential at the bin centres from the bin counts."""
    counts = np.ascontiguousarray(counts, dtype=np.int64)
    wt = np.ascontiguousarray(wake_tab, dtype=np.float64)
    lt = np.ascontiguousarray(lam_tab, dtype=np.float64)
    P = np.empty(counts.size)
    check(_lib.load().hm_wake_potential(_dev(comm).ctx, iptr(counts), counts.size, dptr(wt), dptr(lt),
                                        current, fft_block, dptr(P)))
    return P


def interpolate_kick(particles: Particles, bin_centers: np.ndarray, P: np.ndarray, comm=None):
    """zero_alloc_interpolation! and the kick (src/wakefield.jl:22-105,342-344)."""
    c = particles.coordinates
    ctr = np.ascontiguousarray(bin_centers, dtype=np.float64)
    P = np.ascontiguousarray(P, dtype=np.float64)
    check(_lib.load().hm_interpolate_kick(_dev(comm).ctx, dptr(c.z), dptr(c.ΔE), c.z.size, dptr(ctr),
                                          ctr.size, dptr(P)))


def calculate_wake_function(z, wake_factor, wake_sqrt, inv_cτ):
    """src/wakefield.jl:115-118 (host)."""
    return 0.0 if z > 0 else (wake_factor * math.exp(z * inv_cτ)) * math.cos(wake_sqrt * z)


def delta(x, σ):
    """src/utils.jl:58-66 (host)."""
    return (1 / (math.sqrt(2 * math.pi) * σ)) * math.exp((x * x) * (-0.5 / (σ * σ)))


def apply_wakefield_inplace(particles: Particles, buffers: SimulationBuffers, wake_factor: float,
                            wake_sqrt: float, cτ: float, current: float, sigma_z: float,
                            bin_edges: np.ndarray, bin_centers: np.ndarray, comm=None,
                            use_mpi: bool = True):
    """apply_wakefield_inplace! (src/wakefield.jl:137-345), MPI branch with
    one rank: histogram, smoothed line density, FFT convolution with the wake,
    interpolation and kick, each on the GPU."""
    nbins = buffers.nbins
    if nbins <= 0:
        raise HaissinskiError(_lib.HM_ERR_NBINS, "apply_wakefield_inplace: nbins is zero or negative.")
    if len(bin_edges) != nbins + 1:
        raise HaissinskiError(_lib.HM_ERR_ARG, "apply_wakefield_inplace: bin_edges length mismatch.")
    inv_cτ = math.inf if cτ == 0 else 1 / cτ
    centers = np.asarray(bin_centers, dtype=np.float64)
    counts = histogram(particles.coordinates.z, np.asarray(bin_edges, dtype=np.float64), comm)
    wt = np.array([calculate_wake_function(c, wake_factor, wake_sqrt, inv_cτ) for c in centers])
    lt = np.array([delta(c, 0.15) for c in centers])
    buffers.λ[:] = lt
    P = wake_potential(counts, wt, lt, current, comm)
    interpolate_kick(particles, centers, P, comm)
    return P


def draw_normals(seed: int, turn: int, i0: int, n: int, comm=None) -> np.ndarray:
    """The device QE stream: normals of particles [i0, i0+n) at `turn`."""
    out = np.empty(n)
    check(_lib.load().hm_normals(_dev(comm).ctx, seed & 0xFFFFFFFFFFFFFFFF, turn, i0, n, dptr(out)))
    return out


def compute_global_std(x: np.ndarray, comm=None) -> float:
    """compute_global_std (src/utils.jl:569-606) reduced on the GPU."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = C.c_double()
    check(_lib.load().hm_global_std(_dev(comm).ctx, dptr(x), x.size, C.byref(out)))
    return out.value


def pair_stats(z: np.ndarray, dE: np.ndarray, comm=None) -> dict:
    """Two-pass mean, compute_std (src/utils.jl:531-544) and Statistics.cor of
    (z, ΔE), reduced on the GPU."""
    z = np.ascontiguousarray(z, dtype=np.float64)
    dE = np.ascontiguousarray(dE, dtype=np.float64)
    if z.size != dE.size:
        raise ValueError("z and ΔE differ in length")
    out = np.zeros(5)
    check(_lib.load().hm_pair_stats(_dev(comm).ctx, dptr(z), dptr(dE), z.size, dptr(out)))
    return {"mean_z": out[0], "mean_dE": out[1], "std_z": out[2], "std_dE": out[3],
            "cor": out[4]}


# ---------------------------------------------------------------------------
# figures of merit (src/parameter_sensitivity/types.jl, figures_of_merit.jl)
# ---------------------------------------------------------------------------

class FigureOfMerit:
    """Abstract figure of merit (parameter_sensitivity/types.jl)."""


class EnergySpreadFoM(FigureOfMerit):
    pass


class BunchLengthFoM(FigureOfMerit):
    pass


class EmittanceFoM(FigureOfMerit):
    pass


def compute_fom(fom: FigureOfMerit, particles: Particles, results, comm=None) -> float:
    """compute_fom (src/parameter_sensitivity/figures_of_merit.jl:16-48):
    σ_E, σ_z, or the emittance σ_z σ_E sqrt(1 - cor(z, ΔE)^2), with `results`
    = (σ_E, σ_z, E0) of longitudinal_evolve.  The correlation is reduced on
    the GPU; an EvolveResult of a resident bunch already carries it."""
    σ_E, σ_z, _ = results
    if isinstance(fom, EnergySpreadFoM):
        return σ_E
    if isinstance(fom, BunchLengthFoM):
        return σ_z
    if isinstance(fom, EmittanceFoM):
        if isinstance(results, EvolveResult) and particles is None:
            cor = results.corr_z_dE
            if math.isnan(cor):
                raise ValueError("the result carries no cor(z, ΔE): evolve with "
                                 "TrackingOptions(fom_stats=True) or pass the particles")
        else:
            cor = pair_stats(particles.coordinates.z, particles.coordinates.ΔE, comm)["cor"]
        return σ_z * σ_E * math.sqrt(1 - cor ** 2)
    raise TypeError(f"unknown figure of merit {type(fom).__name__}")


def fd_sensitivity(fom: FigureOfMerit, r_minus: EvolveResult, r_plus: EvolveResult,
                   value: float, rel_step: float) -> float:
    """Central finite difference d(FoM)/d(parameter) from a pair of runs at
    value * (1 -+ rel_step) with common random numbers (same particle and QE
    seeds), as the serving loop's sensitivity pairs produce."""
    f_minus = compute_fom(fom, None, r_minus)
    f_plus = compute_fom(fom, None, r_plus)
    return (f_plus - f_minus) / (2.0 * value * rel_step)


def stream_copy(nbytes: int, iters: int = 20, comm=None) -> tuple[float, float]:
    """(ms per copy, GB/s read+write) of a plain device copy of nbytes."""
    ms = C.c_double()
    gbs = C.c_double()
    check(_lib.load().hm_stream_copy(_dev(comm).ctx, int(nbytes), int(iters), C.byref(ms), C.byref(gbs)))
    return ms.value, gbs.value


def bin_grid(z_width: float, nbins: int):
    """range(-zw, zw, length=nbins+1) and its centres (evolution.jl:273-274),
    correctly rounded.  Host only."""
    edges = np.empty(nbins + 1)
    centers = np.empty(nbins)
    step = C.c_double()
    check(_lib.load().hm_bin_grid(z_width, nbins, dptr(edges), dptr(centers), C.byref(step)))
    return edges, centers, step.value


def generate_particles(μ_z, μ_E, σ_z, σ_E, num_particles, energy, mass, ϕs, freq_rf,
                       seed: int = 1234, comm=None):
    """generate_particles (src/evolution.jl:42-133) on the GPU: the sample
    covariance of min(N, 10000) independent draws defines a MvNormal from which
    the N particles are drawn (N >= 2); N == 1 draws independent normals;
    N <= 0 returns an empty bunch.  Returns (particles, σ_E, σ_z, energy) with
    the bunch's compute_std spreads (src/utils.jl:531-544), or the inputs for
    N == 1, or zeros for N <= 0.  Draws come from the Philox stream of `seed`.
    ``mass``, ``ϕs`` and ``freq_rf`` are accepted and unused, as in the
    reference."""
    if num_particles <= 0:
        return Particles(np.empty(0), np.empty(0)), 0.0, 0.0, energy
    b = DeviceBunch(num_particles, _dev(comm))
    try:
        sE, sz = b.generate_particles(seed, μ_z, μ_E, σ_z, σ_E)
        z, dE = b.download()
    finally:
        b.close()
    return Particles(z, dE), sE, sz, energy


def z_to_ϕ(z_val, rf_factor, ϕs):
    """src/utils.jl:462-464"""
    return -(z_val * rf_factor - ϕs)


def ϕ_to_z(ϕ_val, rf_factor, ϕs):
    """src/utils.jl:471-477"""
    if rf_factor == 0:
        raise ValueError("RF factor is zero, cannot convert phase to position.")
    return (-ϕ_val + ϕs) / rf_factor


def calc_rf_factor(freq_rf, β):
    """src/utils.jl:484-491"""
    if β <= 0:
        return math.inf
    return freq_rf * (2 * math.pi) / (β * SPEED_LIGHT)


def next_power_of_two(n: int) -> int:
    """src/utils.jl:119-122"""
    if n <= 0:
        return 1
    return 1 << (int(n) - 1).bit_length()
