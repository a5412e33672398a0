# IMCHip.jl -- ccall shim that lets MixedPrecisionIMC.jl's main() run its particle loop
# on libimc_hip (MI355X / gfx950).
#
# Usage inside the reference's main loop (src/MixedPrecisionIMC.jl:136-149 in 1D,
# :165-174 in 2D), after mesh generation and Random.seed!:
#
#     eng = IMCHip.Engine(inputs, mesh, simvars)
#     while simvars.t <= simvars.t_end
#         Update.update(inputs, mesh, simvars)
#         IMCHip.sourcing!(eng, mesh, simvars)          # replaces Sourcing.sourcing
#         IMCHip.transport!(eng, mesh, simvars)         # replaces Transport.MC / MC2D / MC_RW
#         IMCHip.clean!(eng)                            # replaces Clean.clean
#         IMCHip.tally!(inputs, mesh, simvars, eng)     # replaces Tally.tally
#         EnergyCheck.energychecker(inputs, mesh, simvars, eng)
#         Output.plotting(inputs, mesh, simvars)
#         timestep(timestepping, simvars)
#     end
#
# or, with the update and the temperature update on the GPU as well (one wait per step):
#
#     IMCHip.set_material!(eng, inputs, mesh, simvars)
#     while simvars.t <= simvars.t_end
#         IMCHip.step!(eng, mesh, simvars)              # replaces update, sourcing, MC, clean, tally
#         EnergyCheck.energychecker(inputs, mesh, simvars, eng)
#         Output.plotting(inputs, mesh, simvars)
#         timestep(timestepping, simvars)
#     end
#
# Arrays cross the boundary as Float64 (column-major, as the reference stores them).
# This file is not exercised in this repository's test suite (no Julia toolchain is
# available where it is built); the same C ABI is exercised through ctypes by tests/.
module IMCHip

const LIB = joinpath(@__DIR__, "..", "csrc", "libimc_hip.so")

const IMC_VACUUM = Int32(0)
const IMC_REFLECT = Int32(1)

struct ImcConfig
    geometry::Int32
    precision::Int32
    nx::Int32
    ny::Int32
    n_scales::Int32
    quirks::Int32
    bc::NTuple{4,Int32}
    seed::UInt32
    device::Int32
    energyscales::NTuple{16,Float64}
    phys_c::Float64
    phys_a::Float64
    distancescale::Float64
    capacity::Int64
    reserved::NTuple{8,Int32}
end

mutable struct Engine
    ptr::Ptr{Cvoid}
    geometry::Int
    nx::Int
    ny::Int
    K::Int
    step::Int
    n::Int64
    function Engine(inputs, mesh, simvars; device::Integer = 0, quirks::Integer = 0)
        geom = simvars.geometry == "1D" ? 1 : 2
        nx, ny = geom == 1 ? (Int(mesh.Ncells), 1) : (Int(mesh.Ncells[1]), Int(mesh.Ncells[2]))
        scales = collect(Float64, mesh.energyscales isa Number ? [mesh.energyscales] : mesh.energyscales)
        prec = Dict(Float16 => 16, Float32 => 32, Float64 => 64)[simvars.precision]
        bcode(s) = uppercase(string(s)) == "REFLECT" ? IMC_REFLECT : IMC_VACUUM
        bc = geom == 1 ? (bcode(simvars.BC[1]), bcode(simvars.BC[2]), IMC_VACUUM, IMC_VACUUM) :
             (bcode(simvars.BC[1]), bcode(simvars.BC[2]), bcode(simvars.BC[3]), bcode(simvars.BC[4]))
        es = ntuple(i -> i <= length(scales) ? scales[i] : 0.0, 16)
        cap = Int64(round(1.25 * (simvars.n_max + simvars.n_input))) + 4 * nx * ny * (Int(simvars.cellmin) + 1)
        cfg = ImcConfig(geom, prec, nx, ny, length(scales), quirks, bc, UInt32(parse(Int, inputs["SEED"])),
                        device, es, Float64(Main.MixedPrecisionIMC.Constants.phys_c),
                        Float64(Main.MixedPrecisionIMC.Constants.phys_a), Float64(mesh.distancescale),
                        cap, ntuple(_ -> Int32(0), 8))
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:imc_create, LIB), Cint, (Ref{Ptr{Cvoid}}, Ref{ImcConfig}), h, cfg)
        eng = new(h[], geom, nx, ny, length(scales), 0, 0)
        check(eng, rc)
        dy = geom == 1 ? C_NULL : Float64.(mesh.dy)
        dx = Float64.(mesh.dx)
        GC.@preserve dx dy check(eng, ccall((:imc_set_geometry, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                                            eng.ptr, dx, dy))
        finalizer(e -> ccall((:imc_destroy, LIB), Cint, (Ptr{Cvoid},), e.ptr), eng)
        return eng
    end
end

Base.length(e::Engine) = Int(e.n)

function check(e::Engine, rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:imc_last_error, LIB), Cstring, (Ptr{Cvoid},), e.ptr))
    error("libimc_hip status $rc: $msg")
end

# Sourcing.sourcing (src/imc_sourcing.jl:12-370)
function sourcing!(e::Engine, mesh, simvars)
    sa = Float64.(vec(e.geometry == 1 ? mesh.sigma_a[:, 1] : mesh.sigma_a[:, :, 1]))
    ss = Float64.(vec(e.geometry == 1 ? mesh.sigma_s[:, 1] : mesh.sigma_s[:, :, 1]))
    fl = Float64.(vec(mesh.fleck))
    te = Float64.(vec(mesh.temp))
    rs = Float64.(vec(mesh.radsource))
    ts = e.geometry == 1 ? [Float64[mesh.temp_surf[1]], Float64[mesh.temp_surf[2]]] :
         [Float64.(vec(mesh.temp_surf[i])) for i in 1:4]
    while length(ts) < 4
        push!(ts, [0.0])
    end
    tptr = [pointer(t) for t in ts]
    GC.@preserve sa ss fl te rs ts tptr begin
        check(e, ccall((:imc_set_step_fields, LIB), Cint,
                       (Ptr{Cvoid}, Int64, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                        Ptr{Float64}, Ptr{Ptr{Float64}}),
                       e.ptr, e.step, Float64(simvars.dt), sa, ss, fl, te, rs, tptr))
    end
    emitted = zeros(Float64, e.nx * e.ny * e.K)
    etot = Ref{Float64}(0.0)
    n = Ref{Int64}(0)
    check(e, ccall((:imc_source, LIB), Cint,
                   (Ptr{Cvoid}, Int64, Int64, Float64, Ptr{Float64}, Ref{Float64}, Ref{Int64}),
                   e.ptr, simvars.n_input, simvars.n_max, Float64(simvars.cellmin), emitted, etot, n))
    mesh.emittedenergy = e.geometry == 1 ? reshape(emitted, e.nx, e.K) : reshape(emitted, e.nx, e.ny, e.K)
    mesh.totalenergy = etot[]
    e.n = n[]
    return nothing
end

# Transport.MC / MC2D (src/imc_transport.jl:13-210, :483-732)
function transport!(e::Engine, mesh, simvars)
    dep = zeros(Float64, e.nx * e.ny * e.K)
    lost = Ref{Float64}(0.0)
    segs = Ref{Int64}(0)
    check(e, ccall((:imc_transport, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ref{Float64}, Ref{Int64}),
                   e.ptr, dep, lost, segs))
    mesh.energydep = e.geometry == 1 ? reshape(dep, e.nx, e.K) : reshape(dep, e.nx, e.ny, e.K)
    mesh.lostenergy += lost[]
    simvars.iterations += segs[]
    return nothing
end

# Clean.clean (src/imc_clean.jl:6-19)
function clean!(e::Engine)
    n = Ref{Int64}(0)
    check(e, ccall((:imc_clean, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), e.ptr, n))
    e.n = n[]
    return nothing
end

# Tally.tally (src/imc_tally.jl:11-143): host part unchanged, census part from the GPU.
function tally!(inputs, mesh, simvars, e::Engine)
    U = Main.MixedPrecisionIMC.Utilities
    C = Main.MixedPrecisionIMC.Constants
    sa = e.geometry == 1 ? mesh.sigma_a[:, 1] : mesh.sigma_a[:, :, 1]
    if simvars.t == 0.0
        for i in eachindex(mesh.temp)
            (mesh.matenergydens[i], _, _) = U.sorter([mesh.fleck[i], sa[i], C.phys_a, C.phys_c, mesh.temp[i],
                                                     mesh.temp[i], mesh.temp[i], mesh.temp[i], simvars.dt,
                                                     mesh.distancescale], 1, Float64)
        end
    end
    mesh.totalenergydep = 0.0
    nrg_inc = zeros(Float64, size(mesh.temp))
    scales = mesh.energyscales isa Number ? [mesh.energyscales] : mesh.energyscales
    for ii in eachindex(scales)
        if e.geometry == 1
            nrg_inc += (mesh.energydep[:, ii] .- mesh.emittedenergy[:, ii]) / scales[ii]
            mesh.totalenergydep += sum((mesh.energydep[:, ii] .* mesh.dx) ./ scales[ii])
        else
            nrg_inc += (mesh.energydep[:, :, ii] .- mesh.emittedenergy[:, :, ii]) / scales[ii]
            mesh.totalenergydep += sum(mesh.energydep[:, :, ii] .* (mesh.dx .* mesh.dy') ./ scales[ii])
        end
    end
    push!(mesh.energyincrease_saved, nrg_inc)
    mesh.matenergydens += nrg_inc
    if inputs["LINEARIZED"] == "TRUE"
        mesh.temp = mesh.matenergydens .^ (1 / 4)
    else
        mesh.temp = mesh.temp + nrg_inc ./ mesh.bee
    end
    rad = zeros(Float64, e.nx * e.ny)
    check(e, ccall((:imc_tally_census, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), e.ptr, rad))
    mesh.radenergydens = e.geometry == 1 ? rad : reshape(rad, e.nx, e.ny)
    push!(mesh.temp_saved, copy(mesh.temp))
    push!(mesh.matenergy_saved, copy(mesh.matenergydens))
    push!(mesh.radenergy_saved, copy(mesh.radenergydens))
    e.step += 1
    return nothing
end

# The device-resident step: one imc_step replaces update!/sourcing!/transport!/clean!/tally!
# (src/MixedPrecisionIMC.jl:137-146); the caller keeps energychecker's final arithmetic
# (its radenergy is returned here) and timestep.
struct ImcMaterial
    sigma_a_coef::Ptr{Float64}
    sigma_a_power::Ptr{Float64}
    sigma_s_coef::Ptr{Float64}
    sigma_s_power::Ptr{Float64}
    bee::Ptr{Float64}
    temp::Ptr{Float64}
    radsource::Ptr{Float64}
    t_surf::NTuple{4,Ptr{Float64}}
    alpha::Float64
    linearized::Int32
    marshak::Int32
end

struct ImcStepResult
    totalenergy::Float64
    totalenergydep::Float64
    radenergy::Float64
    lostenergy::Float64
    histories::Int64
    n_alive::Int64
    segments::Int64
end

function set_material!(e::Engine, inputs, mesh, simvars)
    col(a, k) = Float64.(vec(e.geometry == 1 ? a[:, k] : a[:, :, k]))
    sa, sap, ss, ssp = col(mesh.sigma_a, 2), col(mesh.sigma_a, 3), col(mesh.sigma_s, 2), col(mesh.sigma_s, 3)
    bee = Float64.(vec(mesh.bee .* ones(size(mesh.temp))))
    te = Float64.(vec(mesh.temp))
    rs = Float64.(vec(mesh.radsource))
    ts = e.geometry == 1 ? [Float64[mesh.temp_surf[1]], Float64[mesh.temp_surf[2]], [0.0], [0.0]] :
         [Float64.(vec(mesh.temp_surf[i])) for i in 1:4]
    lin = inputs["LINEARIZED"] == "TRUE"
    marshak = e.geometry == 1 && uppercase(inputs["NAME"]) == "MARSHAK WAVE"
    GC.@preserve sa sap ss ssp bee te rs ts begin
        m = ImcMaterial(pointer(sa), pointer(sap), pointer(ss), pointer(ssp), pointer(bee), pointer(te),
                        pointer(rs), ntuple(i -> pointer(ts[i]), 4),
                        Float64(Main.MixedPrecisionIMC.Constants.alpha), Int32(lin), Int32(marshak))
        check(e, ccall((:imc_set_material, LIB), Cint, (Ptr{Cvoid}, Ref{ImcMaterial}), e.ptr, m))
    end
    return nothing
end

function step!(e::Engine, mesh, simvars)
    C = e.nx * e.ny
    temp, rad, mat, inc = zeros(C), zeros(C), zeros(C), zeros(C)
    r = Ref{ImcStepResult}()
    check(e, ccall((:imc_step, LIB), Cint,
                   (Ptr{Cvoid}, Int64, Float64, Int32, Int64, Int64, Float64, Ptr{Float64}, Ptr{Float64},
                    Ptr{Float64}, Ref{ImcStepResult}),
                   e.ptr, e.step, Float64(simvars.dt), Int32(simvars.t == 0.0), simvars.n_input, simvars.n_max,
                   Float64(simvars.cellmin), temp, rad, mat, r))
    check(e, ccall((:imc_step_energy_increase, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), e.ptr, inc))
    shape(v) = e.geometry == 1 ? v : reshape(v, e.nx, e.ny)
    mesh.temp, mesh.radenergydens, mesh.matenergydens = shape(temp), shape(rad), shape(mat)
    mesh.totalenergy = r[].totalenergy
    mesh.totalenergydep = r[].totalenergydep
    mesh.lostenergy += r[].lostenergy
    simvars.iterations += r[].segments
    push!(mesh.energyincrease_saved, shape(inc))
    push!(mesh.temp_saved, copy(mesh.temp))
    push!(mesh.matenergy_saved, copy(mesh.matenergydens))
    push!(mesh.radenergy_saved, copy(mesh.radenergydens))
    e.n = r[].n_alive
    e.step += 1
    return r[].radenergy
end

"""
    random_walk_enable!(e::Engine, on::Bool = true; r_min = 8.0)

The analytic random walk in the `Transport.MC_RW` slot (imc_transport.jl:212-479, corrected):
in 1D over intervals, in 2D over disks, at least `r_min` collision mean free paths from every face
of the cell (`imc_random_walk_enable`).
"""
function random_walk_enable!(e::Engine, on::Bool = true; r_min::Real = 8.0)
    check(e, ccall((:imc_random_walk_enable, LIB), Cint, (Ptr{Cvoid}, Int32, Float64), e.ptr, Int32(on), Float64(r_min)))
    return nothing
end

"""
    sample_planck(e::Engine, temps::AbstractVector; stream = 0)

Planck frequencies `temps[i] * x_i` (`Sourcing.sample_planck`, imc_sourcing.jl:372-399, with
the temperature factor of its commented-out call site), drawn on the GPU in Float64 and keyed
by (seed, stream, i).
"""
function sample_planck(e::Engine, temps::AbstractVector; stream::Integer = 0)
    t = Vector{Float64}(temps)
    out = similar(t)
    check(e, ccall((:imc_sample_planck, LIB), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, UInt32, Ptr{Float64}),
                   e.ptr, length(t), t, UInt32(stream & 0xffffffff), out))
    return out
end

end # module
