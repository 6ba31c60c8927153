# HaissinskiMI355X.jl -- Julia binding of libhaissinski_mi355x.so.
#
# Adds a `longitudinal_evolve!` method to StochasticHaissinski that runs the
# turn loop on an MI355X: put an `MI355X()` token in the `comm` slot of the
# reference call (src/evolution.jl:206-212); particles, parameters and buffers
# are the reference's own types.  Julia is not installed where this library is
# built, so this file is kept to plain `ccall`s of include/haissinski_mi355x.h.
module HaissinskiMI355X

using StructArrays
import ..StochasticHaissinski: longitudinal_evolve!, SimulationParameters, SimulationBuffers,
                               Particle

const LIB = joinpath(@__DIR__, "..", "libhaissinski_mi355x.so")

# hm_params (field order of SimulationParameters, data_structures.jl:81-97)
struct HmParams
    E0::Float64; mass::Float64; voltage::Float64; harmonic::Int64
    radius::Float64; pipe_radius::Float64; alpha_c::Float64; phi_s::Float64; freq_rf::Float64
    n_turns::Int64
    use_wakefield::Int32; update_eta::Int32; update_E0::Int32; SR_damping::Int32; use_excitation::Int32
end

# hm_options (all zero = the reference)
struct HmOptions
    normals::Ptr{Float64}; seed::UInt64; turn_offset::Int64; wake_table::Ptr{Float64}
    delta_sigma::Float64; bin_half_width_sigmas::Float64; charge::Float64
    n_resonators::Int32; timing::Int32; resonators::Ptr{Float64}; fft_block::Int32; block_size::Int32
    monitor::Ptr{Float64}; serial::Int32; fom_stats::Int32; tile_pairs::Int32; run_pairs::Int32
    bin_whole::Int32
end
HmOptions(; seed::Integer=0, serial::Bool=false) =
    HmOptions(C_NULL, UInt64(seed), 0, C_NULL, 0.0, 0.0, 0.0, 0, 0, C_NULL, 0, 0, C_NULL, Int32(serial),
              0, 0, 0, 0)

struct HmResult
    sigma_E::Float64; sigma_z::Float64; E0::Float64; turns_done::Int64
    drift_skipped::Int32; pad::Int32; ms_total::Float64; ms_pass_avg::Float64; ms_conv_avg::Float64
    pass_launches::Int64; sigma_E0::Float64; sigma_z0::Float64; bin_half_width::Float64
    mean_z::Float64; mean_dE::Float64; corr_z_dE::Float64
end

last_error() = unsafe_string(ccall((:hm_last_error, LIB), Cstring, ()))

mutable struct MI355X
    ctx::Ptr{Cvoid}
    seed::UInt64
    function MI355X(device::Integer=0; seed::Integer=0)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:hm_create, LIB), Int32, (Int32, Ref{Ptr{Cvoid}}), device, h)
        st == 0 || error(last_error())
        d = new(h[], UInt64(seed))
        finalizer(x -> ccall((:hm_destroy, LIB), Int32, (Ptr{Cvoid},), x.ctx), d)
        return d
    end
end

function HmParams(p::SimulationParameters)
    HmParams(Float64(p.E0), Float64(p.mass), Float64(p.voltage), Int64(p.harmonic), Float64(p.radius),
             Float64(p.pipe_radius), Float64(p.α_c), Float64(p.ϕs), Float64(p.freq_rf), Int64(p.n_turns),
             Int32(p.use_wakefield), Int32(p.update_η), Int32(p.update_E0), Int32(p.SR_damping),
             Int32(p.use_excitation))
end

"""
    longitudinal_evolve!(particles, params, buffers, dev::MI355X, use_mpi) -> (σ_E, σ_z, E0)

The reference's turn loop (src/evolution.jl:206-471) on the GPU.  `particles`
is mutated in place; `length(buffers.λ)` gives the bin count.  One GPU holds
the whole bunch, so the MPI branch's collectives are on-device reductions;
`use_mpi = false` selects the serial branch (two-pass `compute_std` spreads,
evolution.jl:245-253,461-467), which, as in the reference, cannot run the wake.
"""
function longitudinal_evolve!(particles::StructArray{Particle{Float64}}, params::SimulationParameters,
                              buffers::SimulationBuffers{Float64}, dev::MI355X, use_mpi::Bool)
    z = particles.coordinates.z
    dE = particles.coordinates.ΔE
    p = Ref(HmParams(params))
    o = Ref(HmOptions(seed=dev.seed, serial=!use_mpi))
    r = Ref{HmResult}()
    st = GC.@preserve z dE ccall((:hm_evolve, LIB), Int32,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Int64, Ref{HmParams}, Ref{HmOptions}, Ref{HmResult}),
        dev.ctx, z, dE, length(z), length(buffers.λ), p, o, r)
    st == 0 || error(last_error())
    return r[].sigma_E, r[].sigma_z, r[].E0
end

export MI355X

end # module
