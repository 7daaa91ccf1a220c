# FFSHip.jl -- Julia binding of libffs_hip.so (the C ABI of include/ffs/ffs.h).
#
# Drop-in for the time-step hot path of FluidFlowSimulation.jl: a simulation
# object whose initialize!/advance_in_time!/writestats/writeoutput call the
# MI355X library instead of the Julia/FFTW step, so run_simulation's driver
# loop (FluidFlowSimulation.jl:31-65), the `global` file and the u1.N/u2.N/u3.N
# padded field files stay as they are.  Every ccall below mirrors one
# declaration of ffs.h; tests/test_julia_shim.py checks the correspondence.
module FFSHip

using DelimitedFiles: writedlm

export FFSSimulation, FFSBatch, ffs_params, initialize!, advance_in_time!, writestats, writespectra, writeoutput,
       run_simulation, footprint

const libffs = get(ENV, "FFS_LIB", joinpath(@__DIR__, "..", "ffs", "libffs_hip.so"))

# ---- ffs_params (field order and types exactly as in ffs.h) ----
mutable struct ffs_params
    size::UInt32
    nx::Int32
    ny::Int32
    nz::Int32
    x_domain::Float64
    y_domain::Float64
    z_domain::Float64
    nu::Float64
    cutoff::Float64
    integrator::Int32
    variable_dt::Int32
    cfl::Float64
    dt::Float64
    start::Int64
    start_time::Float64
    dt_stat::Int32
    dt_spec::Int32
    write_time::Int32
    precision::Int32
    les_model::Int32
    smagorinsky_c::Float64
    filter_width::Float64
    density::Int32
    pr::Float64
    density_gradient::Float64
    z_acceleration::Float64
    reference_density::Float64
    forcing::Int32
    forcing_kf::Float64
    forcing_power::Float64
    forcing_tf::Float64
    forcing_alpha::Float64
    forcing_zf_tanh::Int32
    alvelius_p::Float64
    alvelius_c::Float64
    alvelius_kf::Float64
    hyperviscosity::Int32
    hv_nu::Float64
    hv_m::Int32
    sb_initk::Float64
    sb_endk::Float64
    passive_scalar::Int32
    scalar_pr::Float64
    scalar_gradient::Float64
    scalar_direction::Int32
    equation::Int32
    les_cp::Float64
    les_s_model::Int32
    les_test_width::Float64
    les_cmin::Float64
    scalar_les::Int32
    scalar_les_c::Float64
    density_les::Int32
    density_les_c::Float64
    gravity_direction::Int32
    density_hv::Int32
    density_hv_nu::Float64
    density_hv_m::Int32
    density_sb_initk::Float64
    density_sb_endk::Float64
    ffs_params() = new()
end

const FFS_AB3 = Int32(0)
const FFS_ETD3 = Int32(1)
const FFS_EULER = Int32(2)
const FFS_SPACE_REAL = Int32(0)
const FFS_SPACE_FOURIER = Int32(1)
const FFS_EVENT_OUTPUT = Int32(1)
const FFS_SCALAR_ITERATION = Int32(0)
const FFS_SCALAR_TIME = Int32(1)

function check(rc::Integer, h::Ptr{Cvoid})
    rc < 0 || return rc
    msg = unsafe_string(ccall((:ffs_last_error, libffs), Cstring, (Ptr{Cvoid},), h))
    error("libffs_hip error $rc: $msg")
end

abi_version() = ccall((:ffs_abi_version, libffs), Cint, ())

function default_params()
    p = ffs_params()
    check(ccall((:ffs_params_default, libffs), Cint, (Ref{ffs_params},), p), C_NULL)
    return p
end

# parameters(d) of simulationtypes.jl:242-365 -> ffs_params
function params_from_global(d::AbstractDict)
    p = default_params()
    num(k) = Float64(eval(Meta.parse(d[k])))
    p.nx, p.ny, p.nz = parse(Int32, d[:nx]), parse(Int32, d[:ny]), parse(Int32, d[:nz])
    p.x_domain, p.y_domain, p.z_domain = num(:xDomainSize), num(:yDomainSize), num(:zDomainSize)
    p.nu = num(:kinematicViscosity)
    p.dt_stat, p.dt_spec, p.write_time = parse(Int32, d[:dtStat]), parse(Int32, d[:dtSpec]), parse(Int32, d[:writeTime])
    haskey(d, :start) && (p.start = parse(Int64, d[:start]))
    haskey(d, :startTime) && (p.start_time = parse(Float64, d[:startTime]))
    haskey(d, :cutoff) && (p.cutoff = num(:cutoff))
    vts = get(d, :velocityTimeStep, "ETD3rdO")   # any other value selects ETD3rdO (simulationtypes.jl:308-318)
    p.integrator = vts == "Adams_Bashforth3rdO" ? FFS_AB3 : (vts == "Euller" ? FFS_EULER : FFS_ETD3)
    haskey(d, :variableTimestepFlag) && (p.variable_dt = parse(Bool, d[:variableTimestepFlag]))
    haskey(d, :cfl) && (p.cfl = parse(Float64, d[:cfl]))
    haskey(d, :dt) && (p.dt = num(:dt))
    if haskey(d, :lesModel)   # les_types (les_types.jl:360-404); dynamic models are refused by ffs_create
        codes = Dict("Smagorinsky" => 1, "Vreman" => 2, "Silvis" => 3, "productionViscosity" => 4,
                     "stableNonLinear" => 5, "fakeSmagorinsky" => 6, "SandP" => 7, "dynamicSmagorinsky" => 8,
                     "piomelliSmagorinsky" => 9, "dynamicSandP" => 10, "piomelliSandP" => 11)
        ckey = Dict("Smagorinsky" => :smagorinskyConstant, "Vreman" => :vremanConstant,
                    "Silvis" => :silvisConstant, "productionViscosity" => :productionConstant,
                    "stableNonLinear" => :nlConstant)
        m = d[:lesModel]
        p.les_model = Int32(get(codes, m, -1))
        sandp = m in ("SandP", "dynamicSandP", "piomelliSandP")
        sm = sandp ? d[:sLesModel] : m
        sandp && (p.les_s_model = Int32(get(codes, sm, -1)))
        k = get(ckey, sm, nothing)
        k !== nothing && haskey(d, k) && (p.smagorinsky_c = num(k))
        m == "Silvis" && haskey(d, :silvisPConstant) && (p.les_cp = num(:silvisPConstant))
        m == "SandP" && haskey(d, :pConstant) && (p.les_cp = num(:pConstant))
        dynm = m in ("dynamicSmagorinsky", "piomelliSmagorinsky", "dynamicSandP", "piomelliSandP")
        dynm && haskey(d, :TestFilterWidth) && (p.les_test_width = num(:TestFilterWidth))
        dynm && haskey(d, :cmin) && (p.les_cmin = parse(Float64, d[:cmin]))
        haskey(d, :filterWidth) && (p.filter_width = num(:filterWidth))
    end
    if haskey(d, :densityStratification)
        p.density = Int32(1)
        p.pr, p.density_gradient = num(:Pr), num(:densityGradient)
        p.z_acceleration, p.reference_density = num(:zAcceleration), num(:referenceDensity)
        p.gravity_direction = Int32(Dict("x" => 0, "y" => 1, "z" => 2)[get(d, :gravityDirection, "z")])
    end
    # equation (simulationtypes.jl:265-273): "NS" selects the conservation form
    p.equation = Int32(get(d, :equation, "") == "NS" ? 1 : 0)
    # passive scalar (simulationtypes.jl:320-352)
    if haskey(d, :passiveScalar)
        p.passive_scalar = Int32(1)
        p.scalar_pr, p.scalar_gradient = num(:scalarPr), num(:scalarGradient)
        p.scalar_direction = Int32(Dict("x" => 0, "y" => 1, "z" => 2)[get(d, :scalarDirection, "z")])
    end
    # hyperviscosity (simulationtypes.jl:293-305)
    if haskey(d, :hyperViscosity)
        if d[:hyperViscosity] == "spectralBarrier"
            p.hyperviscosity = Int32(2)
            haskey(d, :initk) && (p.sb_initk = parse(Float64, d[:initk]))
            haskey(d, :endk) && (p.sb_endk = parse(Float64, d[:endk]))
        else
            p.hyperviscosity = Int32(1)
            p.hv_nu = parse(Float64, d[:hyperViscosityCoefficient])
            haskey(d, :hyperViscosityM) && (p.hv_m = parse(Int32, d[:hyperViscosityM]))
        end
    end
    # keys whose features this build does not have: refused rather than ignored
    haskey(d, :specFil) && error("specFil: the filtered statistics and spectra are not built")
    # the scalars' LES fluxes (simulationtypes.jl:334-346, density_stratification_types.jl:176-184)
    if haskey(d, :scalarLesModel) && d[:scalarLesModel] == "vorticityDiffusion"
        p.scalar_les = Int32(2)
        haskey(d, :scalarVorticityCoefficient) && (p.scalar_les_c = parse(Float64, d[:scalarVorticityCoefficient]))
    end
    if haskey(d, :densityLesModel)
        p.density_les = Int32(d[:densityLesModel] == "vorticityDiffusion" ? 2 : 1)
        haskey(d, :densityVorticityCoefficient) && (p.density_les_c = parse(Float64, d[:densityVorticityCoefficient]))
    end
    # the density's own hyperviscosity (density_stratification_types.jl:137-149)
    if haskey(d, :densityHyperViscosity)
        if d[:densityHyperViscosity] == "spectralBarrier"
            p.density_hv = Int32(2)
            haskey(d, :densityInitk) && (p.density_sb_initk = parse(Float64, d[:densityInitk]))
            haskey(d, :densityEndk) && (p.density_sb_endk = parse(Float64, d[:densityEndk]))
        else
            p.density_hv = Int32(1)
            p.density_hv_nu = parse(Float64, d[:densityHyperViscosityCoefficient])
            haskey(d, :densityHyperViscosityM) && (p.density_hv_m = parse(Int32, d[:densityHyperViscosityM]))
        end
    end
    # forcing_model (forcing_types.jl:338-446): the kz = 0-plane schemes
    codes = Dict("rfForcing" => 2, "QfForcing" => 3, "cRfForcing" => 4, "nRfForcing" => 5,
                 "cNRfForcing" => 6, "aForcing" => 7)
    if haskey(d, :forcing) && haskey(codes, d[:forcing])
        f = d[:forcing]
        p.forcing = Int32(codes[f])
        p.forcing_kf = parse(Float64, d[:kf])
        if f == "QfForcing"
            p.alvelius_p, p.alvelius_c = parse(Float64, d[:alveliusP]), parse(Float64, d[:alveliusC])
            p.alvelius_kf = parse(Float64, d[:alveliusKf])
        else
            p.forcing_tf, p.forcing_alpha = parse(Float64, d[:TF]), parse(Float64, d[:alphac])
            zf = get(d, :Zf, "cutoff")   # calculate_Zf, utils.jl:184-218
            p.forcing_zf_tanh = Int32(f == "aForcing" ? zf == "tanh" : zf != "cutoff")
        end
    end
    return p
end

mutable struct FFSSimulation
    h::Ptr{Cvoid}
    p::ffs_params
    function FFSSimulation(p::ffs_params)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:ffs_create, libffs), Cint, (Ref{ffs_params}, Ref{Ptr{Cvoid}}), p, out), C_NULL)
        s = new(out[], p)
        finalizer(x -> ccall((:ffs_destroy, libffs), Cvoid, (Ptr{Cvoid},), x.h), s)
        return s
    end
end

padded_length(s::FFSSimulation) = 2 * (div(s.p.nx, 2) + 1) * s.p.ny * s.p.nz

function set_field!(s::FFSSimulation, field::Integer, a::Array{Float64}; space = FFS_SPACE_REAL)
    GC.@preserve a check(ccall((:ffs_set_field, libffs), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Csize_t),
                               s.h, field, space, a, length(a)), s.h)
end

seed_velocity!(s::FFSSimulation, seed::Integer; k0 = 4.0, energy = 0.5) =
    check(ccall((:ffs_seed_velocity, libffs), Cint, (Ptr{Cvoid}, UInt64, Float64, Float64), s.h, seed, k0, energy), s.h)
seed_density!(s::FFSSimulation, seed::Integer; k0 = 4.0, rms = 0.01) =
    check(ccall((:ffs_seed_density, libffs), Cint, (Ptr{Cvoid}, UInt64, Float64, Float64), s.h, seed, k0, rms), s.h)

function get_field(s::FFSSimulation, field::Integer; space = FFS_SPACE_REAL)
    a = Vector{Float64}(undef, padded_length(s))
    GC.@preserve a check(ccall((:ffs_get_field, libffs), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Csize_t),
                               s.h, field, space, a, length(a)), s.h)
    return a
end

# the real-space read-out without device allocation, into a caller-owned buffer (for callers that
# cannot allocate device memory); serialise calls on one GPU
function get_field_noalloc!(a::Vector{Float64}, s::FFSSimulation, field::Integer)
    GC.@preserve a check(ccall((:ffs_get_field_noalloc, libffs), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Csize_t),
                               s.h, field, a, length(a)), s.h)
    return a
end

initialize!(s::FFSSimulation) = check(ccall((:ffs_initialize, libffs), Cint, (Ptr{Cvoid},), s.h), s.h)

function advance!(s::FFSSimulation, n::Integer)
    done = Ref{Int64}(0)
    ev = Ref{Cint}(0)
    check(ccall((:ffs_advance, libffs), Cint, (Ptr{Cvoid}, Int64, Ref{Int64}, Ref{Cint}), s.h, n, done, ev), s.h)
    return done[], ev[]
end

# advance_in_time!(s) (models.jl:3-17): one step; writes the output files first when due
function advance_in_time!(s::FFSSimulation)
    done, ev = advance!(s, 1)
    if ev == FFS_EVENT_OUTPUT
        writeoutput(s)
        advance!(s, 1)
    end
    writestats(s)
    writespectra(s)
    return nothing
end

# plane forcing state per 2D shell (aForcing: per mode of its (NX, NY) matrices, x fastest):
# what = 0 R, 1 Em, 2 KH (mean |k_h|), 3 Ef (current spectrum), 4 N
function forcing_state(s::FFSSimulation, what::Integer)
    n = Ref{Cint}(0)
    check(ccall((:ffs_forcing_state, libffs), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Cint, Ref{Cint}),
                s.h, what, 0, C_NULL, 0, n), s.h)
    a = zeros(n[])
    check(ccall((:ffs_forcing_state, libffs), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Cint, Ref{Cint}),
                s.h, what, 0, a, length(a), n), s.h)
    return a
end

set_forcing_state!(s::FFSSimulation, what::Integer, a::Vector{Float64}) =
    check(ccall((:ffs_forcing_state, libffs), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Cint, Ref{Cint}),
                s.h, what, 1, a, length(a), Ref{Cint}(0)), s.h)

# exact-resume checkpoints (the reference restarts without its integrator history)
function checkpoint(s::FFSSimulation)
    n = Ref{Csize_t}(0)
    check(ccall((:ffs_checkpoint_size, libffs), Cint, (Ptr{Cvoid}, Ref{Csize_t}), s.h, n), s.h)
    buf = Vector{UInt8}(undef, n[])
    check(ccall((:ffs_checkpoint_save, libffs), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Csize_t), s.h, buf, n[]), s.h)
    return buf
end

restore!(s::FFSSimulation, blob::Vector{UInt8}) =
    check(ccall((:ffs_checkpoint_load, libffs), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Csize_t), s.h, blob, length(blob)), s.h)

writestats_now!(s::FFSSimulation) = check(ccall((:ffs_writestats_now, libffs), Cint, (Ptr{Cvoid},), s.h), s.h)

function stats_header(s::FFSSimulation, table::Integer)
    buf = Vector{UInt8}(undef, 4096)
    check(ccall((:ffs_stats_header, libffs), Cint, (Ptr{Cvoid}, Cint, Ptr{UInt8}, Csize_t), s.h, table, buf, 4096), s.h)
    return unsafe_string(pointer(buf))
end

# writeheader / writestats (stats.jl:9-62): appends every queued record
function writestats(s::FFSSimulation)
    a, b, c = zeros(64), zeros(64), zeros(64)
    na, nb, nc = Ref{Cint}(0), Ref{Cint}(0), Ref{Cint}(0)
    while true
        r = check(ccall((:ffs_stats_pop, libffs), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Cint, Ref{Cint}, Ptr{Float64}, Cint, Ref{Cint}, Ptr{Float64}, Cint,
                         Ref{Cint}), s.h, a, 64, na, b, 64, nb, c, 64, nc), s.h)
        r == 0 && return nothing
        for (name, row, n) in (("Stats.txt", a, na[]), ("Scales.txt", b, nb[]), ("Numbers.txt", c, nc[]))
            open(name, "a+") do f
                join(f, (Int(row[1]), row[2:n]..., "\n"), ",", "")
            end
        end
    end
end

function spectrum(s::FFSSimulation)
    h, v = zeros(8192), zeros(8192)
    nh, nv = Ref{Cint}(0), Ref{Cint}(0)
    check(ccall((:ffs_spectrum, libffs), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint, Ref{Cint}, Ptr{Float64}, Cint, Ref{Cint}),
                s.h, h, 8192, nh, v, 8192, nv), s.h)
    return h[1:2nh[]], v[1:2nv[]]
end

# KH and KRZ (Globals.jl:148-152): the first column of the specH / specV files
function spectra_axes(s::FFSSimulation)
    kh, krz = zeros(8192), zeros(8192)
    nh, nv = Ref{Cint}(0), Ref{Cint}(0)
    check(ccall((:ffs_spectra_axes, libffs), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Ref{Cint}, Ptr{Float64}, Cint, Ref{Cint}),
                s.h, kh, 8192, nh, krz, 8192, nv), s.h)
    return kh[1:nh[]], krz[1:nv[]]
end

const SPEC_QUANTITIES = ("u_h", "u_v", "vis_h", "vis_v", "nl_h", "nl_v", "press_h", "press_v", "les_h", "les_v",
                         "buoyancy_v", "rho", "rhodiss", "rhonl", "forcing_h", "hvis_h", "hvis_v")

# write_spec / write_spec_forcing (spec.jl:1-168): the queued spectra records as the reference's files
function writespectra(s::FFSSimulation)
    kh, krz = spectra_axes(s)
    nh, nv = length(kh), length(krz)
    buf = zeros(length(SPEC_QUANTITIES) * (nh + nv))
    it, n = Ref{Int64}(0), Ref{Csize_t}(0)
    p = s.p
    while true
        r = check(ccall((:ffs_spectra_pop, libffs), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Float64}, Csize_t, Ref{Csize_t}),
                        s.h, it, buf, length(buf), n), s.h)
        r == 0 && return nothing
        i = it[]
        q = Dict(name => (buf[(k - 1) * (nh + nv) + 1:(k - 1) * (nh + nv) + nh],
                          buf[(k - 1) * (nh + nv) + nh + 1:k * (nh + nv)]) for (k, name) in enumerate(SPEC_QUANTITIES))
        for n_ in ("u", "vis", "nl", "press", "les", "hvis")
            (n_ == "les" && p.les_model == 0) && continue
            (n_ == "hvis" && p.hyperviscosity == 0) && continue
            (hH, hV), (vH, vV) = q[n_ * "_h"], q[n_ * "_v"]
            writedlm("$(n_)_h.specH.$i", zip(kh, hH)); writedlm("$(n_)_h.specV.$i", zip(krz, hV))
            writedlm("$(n_).specH.$i", zip(kh, hH .+ vH))
            writedlm("$(n_)_v.specH.$i", zip(kh, vH)); writedlm("$(n_)_v.specV.$i", zip(krz, vV))
        end
        singles = String[]
        if p.density != 0
            bH, bV = q["buoyancy_v"]
            gz = p.z_acceleration / p.reference_density
            writedlm("rhosource.specH.$i", zip(kh, (-).(bH ./ gz) .* p.density_gradient))
            writedlm("rhosource.specV.$i", zip(krz, (-).(bV ./ gz) .* p.density_gradient))
            append!(singles, ["buoyancy_v", "rho", "rhodiss", "rhonl"])
        end
        p.forcing >= 2 && push!(singles, "forcing_h")
        for n_ in singles
            writedlm("$(n_).specH.$i", zip(kh, q[n_][1])); writedlm("$(n_).specV.$i", zip(krz, q[n_][2]))
        end
    end
end

function scalar(s::FFSSimulation, what::Integer)
    out = Ref{Float64}(0)
    check(ccall((:ffs_get_scalar, libffs), Cint, (Ptr{Cvoid}, Cint, Ref{Float64}), s.h, what, out), s.h)
    return out[]
end

profile_enable!(s::FFSSimulation, on::Bool) =
    check(ccall((:ffs_profile_enable, libffs), Cint, (Ptr{Cvoid}, Cint), s.h, on), s.h)
profile_reset!(s::FFSSimulation) = check(ccall((:ffs_profile_reset, libffs), Cint, (Ptr{Cvoid},), s.h), s.h)

function profile(s::FFSSimulation)
    names = Vector{Cstring}(undef, 64)
    ms = zeros(64)
    cnt = zeros(Int64, 64)
    n = Ref{Cint}(0)
    check(ccall((:ffs_profile_get, libffs), Cint, (Ptr{Cvoid}, Ptr{Cstring}, Ptr{Float64}, Ptr{Int64}, Cint, Ref{Cint}),
                s.h, names, ms, cnt, 64, n), s.h)
    return Dict(unsafe_string(names[i]) => (ms[i], cnt[i]) for i in 1:min(n[], 64))
end

function device_bytes(s::FFSSimulation)
    b = Ref{Csize_t}(0)
    check(ccall((:ffs_device_bytes, libffs), Cint, (Ptr{Cvoid}, Ref{Csize_t}), s.h, b), s.h)
    return Int(b[])
end

synchronize(s::FFSSimulation) = check(ccall((:ffs_synchronize, libffs), Cint, (Ptr{Cvoid},), s.h), s.h)

# device bytes a run with these parameters needs (admission control; touches no device)
function footprint(p::ffs_params)
    b = Ref{Csize_t}(0)
    check(ccall((:ffs_footprint, libffs), Cint, (Ref{ffs_params}, Ref{Csize_t}), p, b), C_NULL)
    return Int(b[])
end

function device_memory()
    f, t = Ref{Csize_t}(0), Ref{Csize_t}(0)
    check(ccall((:ffs_device_memory, libffs), Cint, (Ref{Csize_t}, Ref{Csize_t}), f, t), C_NULL)
    return Int(f[]), Int(t[])
end

set_graph!(s::FFSSimulation, on::Bool) = check(ccall((:ffs_set_graph, libffs), Cint, (Ptr{Cvoid}, Cint), s.h, on), s.h)
set_fusion!(s::FFSSimulation, on::Bool) = check(ccall((:ffs_set_fusion, libffs), Cint, (Ptr{Cvoid}, Cint), s.h, on), s.h)

# ---- kernel plans (ffs_plan): the choices an autotuner makes; the tuner itself is ffs/tune.py ----
mutable struct ffs_plan
    x_pair::Cint
    x_waves::Cint
    x_split::Cint
    fusion::Cint
    z_slab::Cint
    reserved::NTuple{3,Cint}
    ffs_plan() = new(0, 0, 0, 0, 0, (0, 0, 0))
end

function get_plan(s::FFSSimulation)
    p = ffs_plan()
    check(ccall((:ffs_get_plan, libffs), Cint, (Ptr{Cvoid}, Ref{ffs_plan}), s.h, p), s.h)
    return p
end
set_plan!(s::FFSSimulation, p::ffs_plan) = check(ccall((:ffs_set_plan, libffs), Cint, (Ptr{Cvoid}, Ref{ffs_plan}), s.h, p), s.h)
function plan_options(s::FFSSimulation)
    p = ffs_plan()
    check(ccall((:ffs_plan_options, libffs), Cint, (Ptr{Cvoid}, Ref{ffs_plan}), s.h, p), s.h)
    return p
end
function device_name()
    buf = Vector{UInt8}(undef, 256)
    check(ccall((:ffs_device_name, libffs), Cint, (Ptr{UInt8}, Csize_t), buf, length(buf)), C_NULL)
    return unsafe_string(pointer(buf))
end

# ---- ensembles: runs advanced in lockstep, one graph launch per steady step ----
function batch_check(rc::Integer, b::Ptr{Cvoid})
    rc < 0 || return rc
    msg = unsafe_string(ccall((:ffs_batch_last_error, libffs), Cstring, (Ptr{Cvoid},), b))
    error("libffs_hip batch error $rc: $msg")
end

mutable struct FFSBatch
    h::Ptr{Cvoid}
    runs::Vector{FFSSimulation}          # kept alive while the batch exists
    function FFSBatch(runs::Vector{FFSSimulation})
        hs = [r.h for r in runs]
        out = Ref{Ptr{Cvoid}}(C_NULL)
        batch_check(ccall((:ffs_batch_create, libffs), Cint, (Ptr{Ptr{Cvoid}}, Cint, Ref{Ptr{Cvoid}}),
                          hs, length(hs), out), C_NULL)
        b = new(out[], runs)
        finalizer(x -> ccall((:ffs_batch_destroy, libffs), Cvoid, (Ptr{Cvoid},), x.h), b)
        return b
    end
end

function advance!(b::FFSBatch, n::Integer)
    done, ev, who = Ref{Int64}(0), Ref{Cint}(0), Ref{Cint}(-1)
    batch_check(ccall((:ffs_batch_advance, libffs), Cint, (Ptr{Cvoid}, Int64, Ref{Int64}, Ref{Cint}, Ref{Cint}),
                      b.h, n, done, ev, who), b.h)
    return done[], ev[], who[]
end

synchronize(b::FFSBatch) = batch_check(ccall((:ffs_batch_synchronize, libffs), Cint, (Ptr{Cvoid},), b.h), b.h)
# true when every run's queued work has finished (does not block)
isidle(b::FFSBatch) = batch_check(ccall((:ffs_batch_query, libffs), Cint, (Ptr{Cvoid},), b.h), b.h) == 1
set_merge!(b::FFSBatch, on::Bool) = batch_check(ccall((:ffs_batch_set_merge, libffs), Cint, (Ptr{Cvoid}, Cint), b.h, on), b.h)

# device ms and steps of the last advance! that took a step; nothing while it still runs
function last_advance_ms(b::FFSBatch)
    ms, n = Ref{Float64}(0), Ref{Int64}(0)
    rc = batch_check(ccall((:ffs_batch_last_advance_ms, libffs), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Int64}),
                           b.h, ms, n), b.h)
    return rc == 1 ? nothing : (ms[], n[])
end

# kernel launches per steady step, merged groups and the runs in them
function launch_info(b::FFSBatch)
    l, g, r = Ref{Cint}(0), Ref{Cint}(0), Ref{Cint}(0)
    batch_check(ccall((:ffs_batch_launch_info, libffs), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cint}, Ref{Cint}),
                      b.h, l, g, r), b.h)
    return Int(l[]), Int(g[]), Int(r[])
end

# writeoutput (output.jl:1-8): u1.$it u2.$it u3.$it in the padded real layout
function writeoutput(s::FFSSimulation; closing::Bool = false)
    it = Int(scalar(s, 0))
    for f in 0:2
        write("u$(f + 1).$it", get_field(s, f))
    end
    s.p.passive_scalar != 0 && write("scalar.$it", get_field(s, 4))
    s.p.density != 0 && write("rho.$it", get_field(s, 3))
    if s.p.density != 0 && s.p.density_les != 0 && s.p.les_model != 0   # the density's LES flux
        for c in 1:3
            write("dflux$c.$it", get_field(s, 9 + c))
        end
    end
    if s.p.les_model != 0   # the stress: real-space product, or real!(tau) at the closing output
        for (name, f) in (("T11", 5), ("T12", 6), ("T13", 7), ("T22", 8), ("T23", 9))
            write("$name.$it", get_field(s, f; space = closing ? 2 : 0))
        end
        s.p.les_model >= 8 && write("cs.$it", get_field(s, 13))    # the dynamic models' coefficient
        s.p.les_model >= 10 && write("cp.$it", get_field(s, 14))   # and SandP's P coefficient
    end
    # the forcing state files (output.jl:44-66): R, N (nRf, cNRf), Eh = (mean |k_h|, Ef) (rf, cRf, nRf, cNRf)
    f = s.p.forcing
    if f == 2 || f >= 4
        R = forcing_state(s, 0)
        f == 7 && (R = reshape(R, :, Int(s.p.ny)))   # aForcing: the (NX, NY) matrix
        writedlm("R.$it", R)
    end
    (f == 5 || f == 6) && writedlm("N.$it", forcing_state(s, 4))
    (f == 2 || f == 4 || f == 5 || f == 6) && writedlm("Eh.$it", zip(forcing_state(s, 2), forcing_state(s, 3)))
end

# parameters(d) (simulationtypes.jl:242-365) over the library: the run of the `global` file's keys with
# its u1.$start..u3.$start (and scalar.$start, rho.$start) loaded
function load_simulation(d::AbstractDict)
    p = params_from_global(d)
    s = FFSSimulation(p)
    for f in 0:2
        a = Vector{Float64}(undef, filesize("u$(f + 1).$(p.start)") ÷ 8)
        read!("u$(f + 1).$(p.start)", a)
        set_field!(s, f, a)
    end
    if p.passive_scalar != 0 && isfile("scalar.$(p.start)")   # zero scalar without one (simulationtypes.jl:324)
        a = Vector{Float64}(undef, filesize("scalar.$(p.start)") ÷ 8)
        read!("scalar.$(p.start)", a)
        set_field!(s, 4, a)
    end
    if p.density != 0 && isfile("rho.$(p.start)")   # zero density without one (density_stratification_types.jl:166)
        a = Vector{Float64}(undef, filesize("rho.$(p.start)") ÷ 8)
        read!("rho.$(p.start)", a)
        set_field!(s, 3, a)
    end
    return s
end

# the Stats/Scales/Numbers headers (writeheader, stats.jl:1-30), written when a run starts at iteration 0
function writeheader(s::FFSSimulation)
    for (name, t) in (("Stats.txt", 0), ("Scales.txt", 1), ("Numbers.txt", 2))
        open(f -> write(f, stats_header(s, t), "\n"), name, "w")
    end
end

# run_simulation() (FluidFlowSimulation.jl:23-65) over the library
function run_simulation(d::AbstractDict)
    s = load_simulation(d)
    p = s.p
    nt = haskey(d, :nt) ? parse(Int, d[:nt]) : typemax(Int) - p.start
    lengthtime = haskey(d, :timeDuration) ? parse(Float64, d[:timeDuration]) : Inf
    p.start == 0 && writeheader(s)
    initialize!(s)
    @assert nt + Int(scalar(s, FFS_SCALAR_ITERATION)) >= p.write_time
    @assert nt + Int(scalar(s, FFS_SCALAR_ITERATION)) >= p.dt_stat
    finalstep = Int(scalar(s, FFS_SCALAR_ITERATION)) + nt
    finaltime = scalar(s, FFS_SCALAR_TIME) + lengthtime
    # the loop of FluidFlowSimulation.jl:41-43: stop at finalstep or once the time reaches finaltime;
    # with a finite duration the time is checked after every step, else the library runs to the next event
    while Int(scalar(s, FFS_SCALAR_ITERATION)) < finalstep && scalar(s, FFS_SCALAR_TIME) < finaltime
        left = isinf(finaltime) ? finalstep - Int(scalar(s, FFS_SCALAR_ITERATION)) : 1
        done, ev = advance!(s, left)
        ev == FFS_EVENT_OUTPUT && writeoutput(s)
        writestats(s)
        writespectra(s)
    end
    # the closing calculate_rhs! writes the output when the last iteration is an output iteration,
    # else run_simulation writes it after the forcing (FluidFlowSimulation.jl:46-61)
    due = mod(Int(scalar(s, 0)), p.write_time) == 0
    due && writeoutput(s)
    writestats_now!(s)
    writestats(s)
    writespectra(s)
    due || writeoutput(s; closing = true)
    p.forcing == 2 && writedlm("R.$(Int(scalar(s, 0)))", forcing_state(s, 0))   # FluidFlowSimulation.jl:59-61
    return s
end

end # module
