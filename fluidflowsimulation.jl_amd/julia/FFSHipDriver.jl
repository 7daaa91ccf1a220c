# FFSHipDriver.jl -- libffs_hip.so behind the reference's own driver loop.
#
# FFSHipSimulation is a subtype of FluidFlowSimulation's AbstractSimulation{EquationType,
# VelocityTimeStepType, PassiveScalarType, DensityStratificationType, LESModelType, ForcingType,
# HyperViscosityType} (simulationtypes.jl:2, macros.jl:35).  Its type parameters are the reference's own
# marker types where a component is absent (NoPassiveScalar, NoDensityStratification, NoLESModel,
# NoForcing, NoHyperViscosity, VorticityEquation) and markers of this file where the library provides
# it, so the reference's traits (hasforcing, hasdensity, haspassivescalar, hasles, ...) answer as they
# would for a reference Simulation of the same case.  Its initialize!, advance_in_time!, calculate_rhs!,
# writestats and writeoutput methods call the library through FFSHip, and the fields the driver reads
# (iteration, time, dtoutput, dtstats, u, forcing, densitystratification.ρ, passivescalar.φ,
# lesmodel.tau) are there, so the reference's run_simulation(s, nt, T) (FluidFlowSimulation.jl:31-65)
# runs the MI355X step without a change to the reference:
#
#     using FluidFlowSimulation
#     include("FFSHip.jl"); include("FFSHipDriver.jl")
#     d = FluidFlowSimulation.readglobal()
#     s = FFSHipDriver.simulation(d)               # in place of parameters(d)
#     nt = haskey(d, :nt) ? parse(Int, d[:nt]) : typemax(Int) - s.iteration[]
#     run_simulation(s, nt, haskey(d, :timeDuration) ? parse(Float64, d[:timeDuration]) : Inf)
#
# What the driver does, step by step, and what answers it here:
#   initialize!(s)             writeheader at iteration 0 (FluidFlowSimulation.jl:68), ffs_initialize
#   advance_in_time!(s)        one step (models.jl:3-17): the output files first when the iteration is an
#                              output iteration (as realspace! writes them inside calculate_rhs!), the
#                              queued stats rows and spectra records after it
#   calculate_rhs!(s)          the closing right-hand side: the output files when the final iteration is an
#                              output iteration (realspace!, models.jl:158); the rest of the closing work is
#                              the library's ffs_writestats_now, run by writestats
#   s.forcing(s)               nothing to do on the host: the forcing is applied inside the library's step
#                              and its closing stats row
#   writestats(s)              ffs_writestats_now (rhs, forcing and the stats row of the final state) and
#                              the rows and spectra records it queues; R.N for rfForcing, which the driver
#                              writes only for a reference RfForcing (FluidFlowSimulation.jl:59-61)
#   real!(s.u) ...             nothing to convert: the state stays on the device
#   writeoutput(s)             the closing output of a final iteration that is not an output iteration
#                              (the stress as real!(tau) of the stored transform, as the reference does)
module FFSHipDriver

using DelimitedFiles: writedlm
using FluidFlowSimulation
import FluidFlowSimulation: initialize!, advance_in_time!, calculate_rhs!, writestats, writeoutput
using ..FFSHip

export FFSHipSimulation, simulation

# The state lives on the device: the driver's real!(field) calls have nothing to convert.
struct FFSFields end
FluidFlowSimulation.real!(::FFSFields) = nothing

# components the library provides (type parameters and the fields the driver reads)
struct FFSEquationNS end                       # the conservation form (equation NS)
struct FFSTimeStep end                         # AB3, ETD3 or Euler, stepped by the library
struct FFSPassiveScalar
    φ::FFSFields
end
struct FFSDensity
    ρ::FFSFields
end
struct FFSLES
    tau::FFSFields
end
struct FFSHyperViscosity end
struct FFSForcing end                          # applied inside the library's step
(::FFSForcing)(s) = nothing

mutable struct FFSHipSimulation{E,V,P,D,L,F,H} <: FluidFlowSimulation.AbstractSimulation{E,V,P,D,L,F,H}
    sim::FFSHip.FFSSimulation
    iteration::Base.RefValue{Int}
    time::Base.RefValue{Float64}
    dtoutput::Int
    dtstats::Int
    u::FFSFields
    passivescalar::P
    densitystratification::D
    lesmodel::L
    forcing::F
    hyperviscosity::H
end

# parameters(d) (simulationtypes.jl:242-365) for the library: the run of the `global` file with its
# initial fields loaded, typed by the components it has
function simulation(d::AbstractDict)
    sim = FFSHip.load_simulation(d)
    p = sim.p
    ps = p.passive_scalar != 0 ? FFSPassiveScalar(FFSFields()) : FluidFlowSimulation.NoPassiveScalar()
    ds = p.density != 0 ? FFSDensity(FFSFields()) : FluidFlowSimulation.NoDensityStratification()
    ls = p.les_model != 0 ? FFSLES(FFSFields()) : FluidFlowSimulation.NoLESModel()
    fs = p.forcing != 0 ? FFSForcing() : FluidFlowSimulation.NoForcing()
    hs = p.hyperviscosity != 0 ? FFSHyperViscosity() : FluidFlowSimulation.NoHyperViscosity()
    E = p.equation != 0 ? FFSEquationNS : FluidFlowSimulation.VorticityEquation
    s = FFSHipSimulation{E,FFSTimeStep,typeof(ps),typeof(ds),typeof(ls),typeof(fs),typeof(hs)}(
        sim, Ref(0), Ref(0.0), Int(p.write_time), Int(p.dt_stat), FFSFields(), ps, ds, ls, fs, hs)
    sync!(s)
    return s
end

# the iteration and time the driver reads, from the library after every call that steps
function sync!(s::FFSHipSimulation)
    s.iteration[] = Int(FFSHip.scalar(s.sim, FFSHip.FFS_SCALAR_ITERATION))
    s.time[] = FFSHip.scalar(s.sim, FFSHip.FFS_SCALAR_TIME)
    return s
end

function initialize!(s::FFSHipSimulation)
    s.iteration[] == 0 && FFSHip.writeheader(s.sim)
    FFSHip.initialize!(s.sim)
    sync!(s)
    return nothing
end

function advance_in_time!(s::FFSHipSimulation)
    FFSHip.advance_in_time!(s.sim)
    sync!(s)
    return nothing
end

function calculate_rhs!(s::FFSHipSimulation)
    mod(s.iteration[], s.dtoutput) == 0 && FFSHip.writeoutput(s.sim)
    return nothing
end

function writestats(s::FFSHipSimulation)
    FFSHip.writestats_now!(s.sim)
    FFSHip.writestats(s.sim)
    FFSHip.writespectra(s.sim)
    s.sim.p.forcing == 2 && writedlm("R.$(s.iteration[])", FFSHip.forcing_state(s.sim, 0))
    return nothing
end

writeoutput(s::FFSHipSimulation) = FFSHip.writeoutput(s.sim; closing = true)

end # module
