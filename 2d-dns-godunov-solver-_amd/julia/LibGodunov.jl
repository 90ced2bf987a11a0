# LibGodunov: the reference's explicit-step entry points on libgodunov (the
# MI355X HIP library of this package), called through its C ABI
# (include/godunov.h) with `ccall`.
#
# The functions keep the reference's names, positional arguments and in-place
# outputs (calcOneStageMUSCLcpuFast.jl:3-6, calcOneStageMUSCL.jl:4-10,
# calcOneStageCUDA.jl:4-23, calcOneStageCUDAx32.jl:5-24, calcDiffTerm.jl:31-33,
# evaluate2d.jl:52-60, utilsFVM2dp.jl:299-301) and take its structs as they
# are: `mesh2d_Int32` (primeObjects.jl:109-135), `fields2d`
# (primeObjects.jl:85-106), `THERMOPHYSICS` (thermo.jl:4-10).  The caller
# defines those types (include the reference's primeObjects.jl and thermo.jl
# before this file); the (nCells, k) column-major arrays go to the library
# without copies or transposes.
#
#   include("primeObjects.jl"); include("thermo.jl")
#   include("/path/to/2d-dns-godunov-solver-_amd/julia/LibGodunov.jl")
#   using .LibGodunov
#   calcOneStageMUSCLcpuFast(cellsThreads, 1.0, dt, t, testMesh, testFields, thermo,
#                            UConsCellsOld, FLUXES, UconsDiffTerm, UconsCellsNew)
#
# The numerics the reference picks by which files it includes (flux,
# limiter, boundary conditions, calcDiffTerm's mu and k) are a `GdParams`,
# set once with `configure!`.
module LibGodunov

export configure!, calcOneStageMUSCLcpuFast, calcOneStageMUSCL, calcOneStageCUDA, calcOneStageCUDAx32,
       calcDiffTerm, updateVariablesSA, cells2nodesSolutionReconstructionWithStencils, GodunovRun, run!,
       drag_gradient, conservative

const libgodunov = get(ENV, "LIBGODUNOV", joinpath(@__DIR__, "..", "libgodunov.so"))

# ---- constants of include/godunov.h
const GD_ORDER_AS_GIVEN, GD_ORDER_RCM, GD_ORDER_PERMUTATION, GD_ORDER_HILBERT = Int32.(0:3)
const GD_SCHEME_REF, GD_SCHEME_NS = Int32(0), Int32(1)
const GD_FP64, GD_FP32_MIXED = Int32(0), Int32(1)
const GD_FLUX_AUSM_PLUS, GD_FLUX_ROE, GD_FLUX_AUSM_PLUS_M2020 = Int32.(0:2)
const GD_LIM_REF_HARMONIC, GD_LIM_VAN_LEER_A, GD_LIM_VAN_ALBADA_A, GD_LIM_MINMOD, GD_LIM_VENKATAKRISHNAN = Int32.(0:4)
const GD_ENERGY_AIR_POLY, GD_ENERGY_CONST = Int32(0), Int32(1)
const GD_BC_NONE, GD_BC_FIXED, GD_BC_EXTRAPOLATE, GD_BC_SLIP_WALL, GD_BC_TGV_MIRROR, GD_BC_ISOTHERMAL_WALL = Int32.(0:5)
const GD_MAX_BC = 16
const GD_ARR_U_OLD, GD_ARR_U_NEW, GD_ARR_U_DIFF, GD_ARR_FLUXES, GD_ARR_DELTA, GD_ARR_U_NODES = Int32.(0:5)
const GD_STAGE_MUSCL_FAST, GD_STAGE_MUSCL = Int32(0), Int32(1)
const GD_DT_FIXED, GD_DT_REF_TAU, GD_DT_LOCAL = Int32.(0:2)
const GD_MAX_KERNELS = 8
const GD_ERR_NONPHYSICAL = Int32(-5)

# ---- structs of include/godunov.h (isbits, so Julia lays them out as C does)
struct GdMeshDesc
    n_cells::Int64; n_nodes::Int64; n_neib_cells::Int32; reserved::Int32
    xNodes::Ptr{Float64}; yNodes::Ptr{Float64}; mesh_connectivity::Ptr{Int32}
    cell_nodes_X::Ptr{Float64}; cell_nodes_Y::Ptr{Float64}; cell_mid_points::Ptr{Float64}
    cell_areas::Ptr{Float64}; HX::Ptr{Float64}
    cell_edges_Nx::Ptr{Float64}; cell_edges_Ny::Ptr{Float64}; cell_edges_length::Ptr{Float64}
    cell_stiffness::Ptr{Int32}; cell_clusters::Ptr{Int32}; node_stencils::Ptr{Float64}
    node2cellsL2up::Ptr{Int32}; node2cellsL2down::Ptr{Int32}; cells2nodes::Ptr{Int32}
end

struct GdBC
    kind::Int32; reserved::Int32; data::NTuple{4,Float64}
end

mutable struct GdParams
    scheme::Int32; precision::Int32; flux::Int32; limiter::Int32
    limiter_eps::Float64; venkat_K::Float64; venkat_scale::NTuple{4,Float64}
    viscous::Int32; energy_law::Int32; mu_const::Float64; k_const::Float64
    RGAS::Float64; Gamma::Float64; mu_ref::Float64; T_ref::Float64; S_suth::Float64; Pr::Float64
    bc::NTuple{17,GdBC}
    GdParams() = new()
end

mutable struct GdRunParams
    dt_policy::Int32; time_kernels::Int32
    dt::Float64; cfl::Float64; betta::Float64; flow_time::Float64
    residuals::Int32; wall_code::Int32
end

mutable struct GdRunStats
    flow_time::Float64; last_dt::Float64; residuals::NTuple{4,Float64}
    wall_force::NTuple{2,Float64}; wall_force_sum::NTuple{2,Float64}
    n_kernels::Int32; nonphysical::Int32
    kernel_names::NTuple{GD_MAX_KERNELS,NTuple{32,UInt8}}
    kernel_ms_total::NTuple{GD_MAX_KERNELS,Float64}
    kernel_launches::NTuple{GD_MAX_KERNELS,Int64}
    kernel_bytes::NTuple{GD_MAX_KERNELS,Float64}
    GdRunStats() = new()
end

struct GdAdjointParams
    n_steps::Int32; checkpoint_every::Int32; dt::Float64; flow_time::Float64
    wall_code::Int32; reserved::Int32; force_weight::NTuple{2,Float64}
    record_budget::Int64
end

mutable struct GdAdjointStats
    J::Float64; force_sum::NTuple{2,Float64}; flow_time::Float64
    recomputed_steps::Int32; n_checkpoints::Int32; storage_bytes::Int64
    ms_forward::Float64; ms_reverse::Float64
    GdAdjointStats() = new()
end

# ---- errors: every entry point returns 0 or a negative code
function last_error()
    b = Vector{UInt8}(undef, 1024)
    ccall((:gd_last_error, libgodunov), Cint, (Ptr{UInt8}, Csize_t), b, length(b))
    unsafe_string(pointer(b))
end
gdcheck(rc) = rc == 0 ? nothing : error("libgodunov error $rc: ", last_error())

# ---- handles
mutable struct GdMesh
    h::Ptr{Cvoid}
    n_cells::Int
    n_nodes::Int
end

function GdMesh(m, order::Integer = GD_ORDER_HILBERT)
    d = GdMeshDesc(m.nCells, m.nNodes, m.nNeibCells, 0, pointer(m.xNodes), pointer(m.yNodes),
                   pointer(m.mesh_connectivity), pointer(m.cell_nodes_X), pointer(m.cell_nodes_Y),
                   pointer(m.cell_mid_points), pointer(m.cell_areas), pointer(m.HX), pointer(m.cell_edges_Nx),
                   pointer(m.cell_edges_Ny), pointer(m.cell_edges_length), pointer(m.cell_stiffness),
                   pointer(m.cell_clusters), pointer(m.node_stencils), pointer(m.node2cellsL2up),
                   pointer(m.node2cellsL2down), pointer(m.cells2nodes))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve m gdcheck(ccall((:gd_mesh_create, libgodunov), Cint,
                                 (Ref{GdMeshDesc}, Int32, Ptr{Int64}, Ref{Ptr{Cvoid}}), d, order, C_NULL, h))
    gm = GdMesh(h[], m.nCells, m.nNodes)
    finalizer(x -> ccall((:gd_mesh_destroy, libgodunov), Cint, (Ptr{Cvoid},), x.h), gm)
end

mutable struct GdState
    h::Ptr{Cvoid}
    mesh::GdMesh
end

function GdState(gm::GdMesh, p::GdParams)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    gdcheck(ccall((:gd_state_create, libgodunov), Cint, (Ptr{Cvoid}, Ref{GdParams}, Ref{Ptr{Cvoid}}), gm.h, p, h))
    st = GdState(h[], gm)
    finalizer(x -> ccall((:gd_state_destroy, libgodunov), Cint, (Ptr{Cvoid},), x.h), st)
end

set_params!(st::GdState, p::GdParams) =
    gdcheck(ccall((:gd_state_set_params, libgodunov), Cint, (Ptr{Cvoid}, Ref{GdParams}), st.h, p))
set_array!(st::GdState, which, a::Matrix{Float64}) =
    gdcheck(ccall((:gd_state_set_array, libgodunov), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}), st.h, which, a))
get_array!(st::GdState, which, a::Matrix{Float64}) =
    gdcheck(ccall((:gd_state_get_array, libgodunov), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}), st.h, which, a))

function set_fields!(st::GdState, f; nodes::Bool = true)
    N(x) = nodes ? x : Ptr{Float64}(C_NULL)
    gdcheck(ccall((:gd_state_set_fields, libgodunov), Cint,
                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                   Ptr{Float64}, Ptr{Float64}), st.h, f.densityCells, f.UxCells, f.UyCells, f.pressureCells,
                  N(f.densityNodes), N(f.UxNodes), N(f.UyNodes), N(f.pressureNodes)))
end

function get_fields!(st::GdState, f)
    gdcheck(ccall((:gd_state_get_fields, libgodunov), Cint,
                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                   Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), st.h, f.densityCells, f.UxCells,
                  f.UyCells, f.pressureCells, f.aSoundCells, f.VMAXCells, f.densityNodes, f.UxNodes, f.UyNodes,
                  f.pressureNodes))
end

# ---- numerics: the reference's production choice by default (AUSM+, the
# harmonic Minmod_Limiter with 1e-12, boundaryConditions2d.jl's codes -1
# fixed M = 3.55 inflow, -2 zero gradient, -3 slip wall, calcDiffTerm's mu =
# 1e-2 with the air k(T))
function default_params()
    p = GdParams()
    ccall((:gd_params_default, libgodunov), Cvoid, (Ref{GdParams},), p)
    p
end

function bc_table(bcs::Dict)
    t = [GdBC(GD_BC_EXTRAPOLATE, 0, (0.0, 0.0, 0.0, 0.0)) for _ in 0:GD_MAX_BC]
    for (code, (kind, data)) in bcs
        d = ntuple(q -> q <= length(data) ? Float64(data[q]) : 0.0, 4)
        t[1 - code] = GdBC(Int32(kind), 0, d)
    end
    Tuple(t)
end

const REFERENCE_BCS = Dict(-1 => (GD_BC_FIXED, (1.1766, 1232.6445, 0.0, 101325.0)),
                           -2 => (GD_BC_EXTRAPOLATE, ()), -3 => (GD_BC_SLIP_WALL, ()))

const CONFIG = Ref{Any}(nothing)

"""configure!(; flux, limiter, limiter_eps, bcs, energy_law, mu_const, k_const):
the reference scheme's numerics for every later call."""
function configure!(; flux = GD_FLUX_AUSM_PLUS, limiter = GD_LIM_REF_HARMONIC, limiter_eps = 1e-12,
                    bcs = REFERENCE_BCS, energy_law = GD_ENERGY_AIR_POLY, mu_const = 1e-2, k_const = 0.0)
    CONFIG[] = (; flux, limiter, limiter_eps, bcs, energy_law, mu_const, k_const)
end

function ref_params(thermo)
    c = CONFIG[] === nothing ? configure!() : CONFIG[]
    p = default_params()
    p.scheme = GD_SCHEME_REF; p.precision = GD_FP64; p.flux = c.flux; p.limiter = c.limiter
    p.limiter_eps = c.limiter_eps; p.viscous = 1; p.energy_law = c.energy_law
    p.mu_const = c.mu_const; p.k_const = c.k_const
    p.RGAS = thermo.RGAS; p.Gamma = thermo.Gamma
    p.bc = bc_table(c.bcs)
    p
end

# one device mesh and state per reference mesh, created on first use
const CONTEXT = IdDict{Any,Any}()     # mesh => (GdMesh, GdState, the state's GdParams)
const BY_NCELLS = Dict{Int,Any}()     # the last context of each cell count, for updateVariablesSA
function state_for(mesh, thermo)
    p = ref_params(thermo)
    ctx = get!(() -> (gm = GdMesh(mesh); (gm, GdState(gm, p), p)), CONTEXT, mesh)
    set_params!(ctx[2], p)
    CONTEXT[mesh] = BY_NCELLS[mesh.nCells] = (ctx[1], ctx[2], p)
    ctx[2]
end

# cells a partition table covers (partMesh2d.jl: rows of 1-based inclusive
# ranges; the one-thread table of zeros indexes cell 0 and throws there too)
function covered(threads::AbstractMatrix, n::Integer)
    mask = falses(n)
    for p in 1:size(threads, 1)
        b, e = threads[p, 1], threads[p, 2]
        (b < 1 || e > n) && throw(BoundsError(1:n, b:e))
        mask[b:e] .= true
    end
    mask
end

function copy_rows!(dst::AbstractMatrix, src::AbstractMatrix, mask)
    all(mask) ? copyto!(dst, src) : (dst[mask, :] .= src[mask, :])
    dst
end

# ---- the reference's operators

"calcOneStageMUSCLcpuFast.jl:3-107: FLUXES and UconsCellsNew overwritten (with the Udiff term)."
function calcOneStageMUSCLcpuFast(cellsThreads, bettaKJ, dt, flowTime, testMesh, testFields, thermo,
                                  UConsCellsOld, FLUXES, UconsDiffTerm, UconsCellsNew)
    mask = covered(cellsThreads, testMesh.nCells)
    st = state_for(testMesh, thermo)
    set_fields!(st, testFields)
    set_array!(st, GD_ARR_U_OLD, UConsCellsOld)
    set_array!(st, GD_ARR_U_DIFF, UconsDiffTerm)
    gdcheck(ccall((:gd_calc_one_stage, libgodunov), Cint, (Ptr{Cvoid}, Int32, Float64, Float64, Float64),
                  st.h, GD_STAGE_MUSCL_FAST, bettaKJ, dt, flowTime))
    n = testMesh.nCells
    tmp = Matrix{Float64}(undef, n, 4)
    get_array!(st, GD_ARR_FLUXES, tmp); copy_rows!(FLUXES, tmp, mask)
    get_array!(st, GD_ARR_U_NEW, tmp); copy_rows!(UconsCellsNew, tmp, mask)
    nothing
end

# the face states ULeft1..4 / URight1..4 as phase one fills them
function face_states!(st::GdState, flowTime, lefts, rights, mask)
    n = st.mesh.n_cells
    L = Array{Float64}(undef, 4, n, 4); R = Array{Float64}(undef, 4, n, 4)
    for k in 1:4
        lefts[k] !== nothing && (L[:, :, k] .= lefts[k])
        rights[k] !== nothing && (R[:, :, k] .= rights[k])
    end
    gdcheck(ccall((:gd_calc_face_states, libgodunov), Cint, (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}),
                  st.h, flowTime, L, R))
    for k in 1:4
        lefts[k] !== nothing && (lefts[k][:, mask] .= L[:, mask, k])
        rights[k] !== nothing && (rights[k][:, mask] .= R[:, mask, k])
    end
end

"calcOneStageMUSCL.jl:4-106: two-phase stage without the Udiff term; fills ULeft*/URight* and FLUXES."
function calcOneStageMUSCL(betta, dtX, flowTime, testMesh, testfields2dX, thermo, cellsThreadsX,
                           UconsCellsOldX, UconsDiffTermX, UconsCellsNewX, ULeft1, ULeft2, ULeft3, ULeft4,
                           URight1, URight2, URight3, URight4, FLUXES)
    mask = covered(cellsThreadsX, testMesh.nCells)
    st = state_for(testMesh, thermo)
    set_fields!(st, testfields2dX)
    set_array!(st, GD_ARR_U_OLD, UconsCellsOldX)
    gdcheck(ccall((:gd_calc_one_stage, libgodunov), Cint, (Ptr{Cvoid}, Int32, Float64, Float64, Float64),
                  st.h, GD_STAGE_MUSCL, betta, dtX, flowTime))
    n = testMesh.nCells
    tmp = Matrix{Float64}(undef, n, 4)
    get_array!(st, GD_ARR_U_NEW, tmp); copy_rows!(UconsCellsNewX, tmp, mask)
    get_array!(st, GD_ARR_FLUXES, tmp); copy_rows!(FLUXES, tmp, mask)
    face_states!(st, flowTime, (ULeft1, ULeft2, ULeft3, ULeft4), (URight1, URight2, URight3, URight4), mask)
    nothing
end

function cuda_stage(precision, betta, dtX, flowTime, testMeshX, thermoX, cellsThreadsX, UconsCellsOldX,
                    UconsCellsNewX, lefts, rights, fluxes)
    mask = covered(cellsThreadsX, testMeshX.nCells)
    st = state_for(testMeshX, thermoX)
    set_array!(st, GD_ARR_U_OLD, UconsCellsOldX)
    n = testMeshX.nCells
    L = Array{Float64}(undef, 4, n, 4); R = Array{Float64}(undef, 4, n, 4)
    for k in 1:3
        L[:, :, k] .= lefts[k]; R[:, :, k] .= rights[k]
    end
    iflux = Matrix{Float64}(undef, n, 4)
    gdcheck(ccall((:gd_calc_one_stage_cuda, libgodunov), Cint,
                  (Ptr{Cvoid}, Int32, Float64, Float64, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                  st.h, precision, betta, dtX, flowTime, L, R, iflux))
    tmp = Matrix{Float64}(undef, n, 4)
    get_array!(st, GD_ARR_U_NEW, tmp); copy_rows!(UconsCellsNewX, tmp, mask)
    for k in 1:4
        fluxes[k] .= @view iflux[:, k]   # copyto!(iFluxV_k, cuFluxV_k) copies every cell
    end
    nothing
end

"calcOneStageCUDA.jl:4-135: AUSM+ of the given face states on the GPU (FP64), the update without Udiff.
The CuVector work buffers of the source are accepted and not used."
calcOneStageCUDA(betta, dtX, flowTime, testMeshX, testfields2dX, thermoX, cellsThreadsX, UconsCellsOldX,
                 UconsDiffTermX, UconsCellsNewX, uLeft1, uLeft2, uLeft3, uLeft4, uRight1, uRight2, uRight3, uRight4,
                 iFluxV1, iFluxV2, iFluxV3, iFluxV4, device_buffers...) =
    cuda_stage(GD_FP64, betta, dtX, flowTime, testMeshX, thermoX, cellsThreadsX, UconsCellsOldX, UconsCellsNewX,
               (uLeft1, uLeft2, uLeft3, uLeft4), (uRight1, uRight2, uRight3, uRight4),
               (iFluxV1, iFluxV2, iFluxV3, iFluxV4))

"calcOneStageCUDAx32.jl:5-155: the same with Float32 face states, kernel arithmetic and per-cell sums."
calcOneStageCUDAx32(betta, dtX, flowTime, testMeshX, testfields2dX, thermoX, cellsThreadsX, UconsCellsOldX,
                    UconsDiffTermX, UconsCellsNewX, uLeft1, uLeft2, uLeft3, uLeft4, uRight1, uRight2, uRight3,
                    uRight4, iFluxV1, iFluxV2, iFluxV3, iFluxV4, device_buffers...) =
    cuda_stage(GD_FP32_MIXED, betta, dtX, flowTime, testMeshX, thermoX, cellsThreadsX, UconsCellsOldX,
               UconsCellsNewX, (uLeft1, uLeft2, uLeft3, uLeft4), (uRight1, uRight2, uRight3, uRight4),
               (iFluxV1, iFluxV2, iFluxV3, iFluxV4))

"calcDiffTerm.jl:31-121: columns 2:4 of UConsDiffCells overwritten (column 1 left as it is)."
function calcDiffTerm(cellsThreads, nodesThreads, testMesh, testfields2d, viscousFields2d, thermo,
                      UCCells, UCNodes, UConsDiffCells)
    mask = covered(cellsThreads, testMesh.nCells)
    st = state_for(testMesh, thermo)
    set_fields!(st, testfields2d; nodes = false)
    set_array!(st, GD_ARR_U_OLD, UCCells)
    set_array!(st, GD_ARR_U_DIFF, UConsDiffCells)
    gdcheck(ccall((:gd_calc_diff_term, libgodunov), Cint, (Ptr{Cvoid},), st.h))
    tmp = Matrix{Float64}(undef, testMesh.nCells, 4)
    get_array!(st, GD_ARR_U_DIFF, tmp)
    copy_rows!(UConsDiffCells, tmp, mask)
    nothing
end

"""evaluate2d.jl:52-110 on the mesh of the last step with as many cells:
primitives of UconsCellsNew into testfields2d, Delta = new - old,
UconsCellsOld = new (dynControls' density range as the source sets it)."""
function updateVariablesSA(cellsThreadsX, Gamma, UconsCellsNew, UconsCellsOld, Delta, testfields2d,
                           solControls = nothing, dynControls = nothing)
    n = size(UconsCellsNew, 1)
    mask = covered(cellsThreadsX, n)
    haskey(BY_NCELLS, n) || error("updateVariablesSA: no step has run on a mesh of $n cells yet")
    _, st, p = BY_NCELLS[n]
    p.Gamma = Gamma
    set_params!(st, p)
    set_array!(st, GD_ARR_U_NEW, UconsCellsNew)
    set_array!(st, GD_ARR_U_OLD, UconsCellsOld)
    gdcheck(ccall((:gd_update_variables, libgodunov), Cint, (Ptr{Cvoid},), st.h))
    tmp = Matrix{Float64}(undef, n, 4)
    get_array!(st, GD_ARR_DELTA, tmp); copy_rows!(Delta, tmp, mask)
    get_array!(st, GD_ARR_U_OLD, tmp); copy_rows!(UconsCellsOld, tmp, mask)
    get_fields!(st, testfields2d)
    if dynControls !== nothing
        dynControls.rhoMax = maximum(testfields2d.densityCells)
        dynControls.rhoMin = minimum(testfields2d.densityCells)
    end
    nothing
end

"utilsFVM2dp.jl:299-365: node primitives into testfields2d, node conservative values into node_solution."
function cells2nodesSolutionReconstructionWithStencils(nodesThreads, testMesh, testfields2d, viscfields2d,
                                                       cell_solution, node_solution; thermo)
    covered(nodesThreads, testMesh.nNodes)
    st = state_for(testMesh, thermo)
    set_fields!(st, testfields2d; nodes = false)
    set_array!(st, GD_ARR_U_OLD, cell_solution)
    gdcheck(ccall((:gd_cells2nodes, libgodunov), Cint, (Ptr{Cvoid},), st.h))
    get_fields!(st, testfields2d)
    get_array!(st, GD_ARR_U_NODES, node_solution)
    nothing
end

# ---- a device-resident run: the step and what the reference's driver does
# around it (updateVariablesSA, node interpolation, dt, residuals, wall force)
# without a host round trip per step

mutable struct GodunovRun
    state::GdState
    flow_time::Float64
end

"GodunovRun(mesh, params; order): a state resident in HBM (params: a GdParams, e.g. for the ns scheme)."
GodunovRun(mesh, p::GdParams; order = GD_ORDER_HILBERT) = GodunovRun(GdState(GdMesh(mesh, order), p), 0.0)

function conservative(r::GodunovRun)
    U = Matrix{Float64}(undef, r.state.mesh.n_cells, 4)
    get_array!(r.state, GD_ARR_U_OLD, U)
    U
end
set_conservative!(r::GodunovRun, U::Matrix{Float64}) = set_array!(r.state, GD_ARR_U_OLD, U)

"run!(r, n; dt, cfl, wall_code): n steps (dt given: fixed; else the local CFL step); returns GdRunStats."
function run!(r::GodunovRun, n::Integer; dt = nothing, cfl = 0.4, wall_code = 0, residuals = false)
    rp = GdRunParams(dt === nothing ? GD_DT_LOCAL : GD_DT_FIXED, 0, dt === nothing ? 0.0 : dt, cfl, 1.0,
                     r.flow_time, residuals, wall_code)
    s = GdRunStats()
    gdcheck(ccall((:gd_run, libgodunov), Cint, (Ptr{Cvoid}, Int32, Ref{GdRunParams}, Ref{GdRunStats}),
                  r.state.h, n, rp, s))
    r.flow_time = rp.flow_time
    s
end

"""drag_gradient(r, N, dt; every, wall, w, dJdUN, budget): the discrete adjoint
(gd_adjoint_run) of J = sum over N fixed-dt steps of w . F_wall (+ <dJdUN, U_N>)
from the run's current state, which ends at U_N.  Returns (dJ/dU0 (n_cells,
4), dJ/dbc (4, 17): column 1 - code for a fixed-state code, stats).  On a
reference-scheme run the wall force is the AUSM+ momentum flux through the
faces of code `wall` (AUSM+ with an odd, symmetric limiter, constant mu and
k)."""
function drag_gradient(r::GodunovRun, N::Integer, dt::Real; every = 20, wall = -3, w = (1.0, 0.0),
                       dJdUN = nothing, budget = 0)
    n = r.state.mesh.n_cells
    g0 = Matrix{Float64}(undef, n, 4)
    gbc = zeros(4, GD_MAX_BC + 1)
    ap = GdAdjointParams(N, every, dt, r.flow_time, wall, 0, (Float64(w[1]), Float64(w[2])), budget)
    s = GdAdjointStats()
    gdcheck(ccall((:gd_adjoint_run, libgodunov), Cint,
                  (Ptr{Cvoid}, Ref{GdAdjointParams}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{GdAdjointStats}),
                  r.state.h, ap, dJdUN === nothing ? Ptr{Float64}(C_NULL) : dJdUN, g0, gbc, s))
    r.flow_time = s.flow_time
    g0, gbc, s
end

end  # module
