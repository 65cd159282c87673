#=
FLOWPanelHIP -- the `ccall` layer that binds FLOWPanel.jl's hot path to the
gfx950 library libflowpanel_hip.so (C ABI: include/flowpanel_hip.h).

The reference's bodies keep their Julia types; a device handle rides along in
an `IdDict` keyed by the body, so the reference's structs stay unchanged.
Layouts are already the reference's: `nodes`, `Uinfs`, `Das`, `Dbs`,
`targets`, `out` are column-major 3 x n Float64, panels 1-based
(`index_base = 1`), `shedding` the body's 6 x ns Int matrix (-1 sentinel).

Entry points mirrored (reference file:line):
  solve(body, Uinfs[, Das, Dbs]; solver)   src/FLOWPanel_liftingbody.jl:172-216, 241-316,
                                            src/FLOWPanel_nonliftingbody.jl:93-130, 258-282
  Uind!(body, targets, out)                src/FLOWPanel_abstractbody.jl:101-110
  phi!(body, targets, out)                 src/FLOWPanel_abstractbody.jl:118-127
  calcfield_U!(out, source, target)        src/FLOWPanel_postprocess.jl:28-99
  calcfield_Cp!(out, body, Us, Uref)       src/FLOWPanel_postprocess.jl:1059-1112
  calcfield_Ugradmu!(out, body)            src/FLOWPanel_postprocess.jl:756-1000
  solve_ludiv!/solve_backslash!/solve_gmres!(y, A, b)
                                           src/FLOWPanel_linearsolver.jl:24-157
plus the library's additions: mixed-precision solvers, many freestreams as
right-hand sides, batches of bodies, and the lift/drag adjoint as a
ChainRulesCore rrule.
=#
module FLOWPanelHIP

import FLOWPanel as pnl
import ChainRulesCore

const lib = joinpath(@__DIR__, "..", "libflowpanel_hip.so")

# ---- status, context ---------------------------------------------------------
const FP_OK = Cint(0)
const SOLVER_LU = Cint(0)
const SOLVER_LU32_IR = Cint(1)
const SOLVER_GMRES_IR = Cint(2)
const SOLVER_GMRES = Cint(3)
const KUTTA_NONE = Cint(0)
const KUTTA_REFERENCE = Cint(1)
const KUTTA_PAIR = Cint(2)

mutable struct Ctx
    h::Ptr{Cvoid}
end
const CTX = Ref{Ctx}()

function ctx()
    if !isassigned(CTX)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:fp_ctx_create, lib), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, r)
        st == FP_OK || error("fp_ctx_create failed ($st): no usable GPU")
        c = Ctx(r[])
        finalizer(x -> ccall((:fp_ctx_destroy, lib), Cvoid, (Ptr{Cvoid},), x.h), c)
        CTX[] = c
    end
    return CTX[].h
end

function check(st)
    st == FP_OK && return nothing
    msg = unsafe_string(ccall((:fp_ctx_last_error, lib), Cstring, (Ptr{Cvoid},), ctx()))
    error("flowpanel_hip error $st: $msg")
end

# ---- structs of the C ABI (field order and types as in the header) ----------
struct SolveOpts
    solver::Cint
    tol::Cdouble
    inner_tol::Cdouble
    restart::Cint
    max_iters::Cint
    max_refine::Cint
    precond_block::Cint
    coarse_block::Cint
    residual_mode::Cint
    precond_order::Cint
end

struct SolveStats
    refinements::Cint
    iterations::Cint
    residual::Cdouble
    ms_assemble::Cdouble
    ms_factor::Cdouble
    ms_solve::Cdouble
    info::Cint
end

struct LiftDragOpts
    probe_offset::Cdouble
    probe_length::Cdouble
    probe_charlen::Cint
    kutta_mode::Cint
    span_dir::NTuple{3,Cdouble}
    ref_area::Cdouble
    solve::SolveOpts
end

struct LiftDragStats
    CL::Cdouble
    CDi::Cdouble
    ms_forward::Cdouble
    ms_adjoint::Cdouble
    ms_transpose_solve::Cdouble
    iterations_adjoint::Cint
    forward::SolveStats
end

struct BatchStats
    ms_graph::Cdouble
    ms_capture::Cdouble
    ms_extra::Cdouble
    residual::Cdouble
    panels::Int64
    nbodies::Cint
    sweeps_in_graph::Cint
    extra_sweeps::Cint
    graph_nodes::Cint
    recaptured::Cint
    tile_targets::Cint
    queue_workgroups::Cint
end

function solve_opts(solver::Integer; kw...)
    o = Ref{SolveOpts}()
    ccall((:fp_solve_opts_default, lib), Cvoid, (Ref{SolveOpts}, Cint), o, solver)
    x = o[]
    vals = Any[getfield(x, f) for f in fieldnames(SolveOpts)]
    for (k, v) in kw
        i = findfirst(==(k), fieldnames(SolveOpts))
        i === nothing && error("unknown solver option $k")
        vals[i] = convert(fieldtype(SolveOpts, i), v)
    end
    return SolveOpts(vals...)
end

# ---- bodies -------------------------------------------------------------------
# the reference's characteristic-length functions (src/FLOWPanel_abstractbody.jl:571-633)
function charlen_kind(f)
    f === pnl.characteristiclength_unitary && return Cint(0)
    f === pnl.characteristiclength_bbox && return Cint(1)
    f === pnl.characteristiclength_maxdist && return Cint(2)
    f === pnl.characteristiclength_sqrtarea && return Cint(3)
    error("characteristic length $f has no device counterpart")
end

element_code(body) = body isa pnl.NonLiftingBody{pnl.ConstantSource} ? Cint(0) : Cint(1)

const HANDLES = IdDict{Any,Ptr{Cvoid}}()

# the body's panels in its own order (get_cell_t!), flattened into 1-based CSR
function handle(body::pnl.AbstractBody)
    get!(HANDLES, body) do
        nodes = body.grid._nodes
        cells = [collect(pnl.gt.get_cell_t(body.grid, i)) for i in 1:body.ncells]
        ptr = Int64.(cumsum(vcat(1, length.(cells))))
        idx = Int64.(reduce(vcat, cells))
        sh = body isa pnl.NonLiftingBody ? zeros(Int64, 6, 0) : Int64.(body.shedding)
        clen = hasproperty(body, :characteristiclength) ? charlen_kind(body.characteristiclength) : Cint(0)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:fp_body_create, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Cint, Cint, Ptr{Int64}, Int64,
                     Cdouble, Cint, Cdouble, Cdouble, Ref{Ptr{Cvoid}}),
                    ctx(), nodes, size(nodes, 2), ptr, idx, body.ncells, 1, element_code(body), sh, size(sh, 2),
                    body.CPoffset, clen, body.kerneloffset, body.kernelcutoff, h))
        h[]
    end
end

"""release!(body): free the device copy of a body (a later call re-creates it)."""
function release!(body)
    for b in components(body)
        h = pop!(HANDLES, b, C_NULL)
        h == C_NULL || ccall((:fp_body_destroy, lib), Cvoid, (Ptr{Cvoid},), h)
    end
end

components(body) = body isa pnl.MultiBody ? body.bodies : [body]
handles(body) = Ptr{Cvoid}[handle(b) for b in components(body)]

function set_freestream!(body, Uinfs, Das, Dbs)
    col = 0
    sh = 0
    for b in components(body)
        U = Uinfs[:, col+1:col+b.ncells]
        ns = b isa pnl.NonLiftingBody ? 0 : b.nsheddings
        da = ns > 0 ? Das[:, sh+1:sh+ns] : C_NULL
        db = ns > 0 ? Dbs[:, sh+1:sh+ns] : C_NULL
        check(ccall((:fp_body_set_freestream, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), ctx(), handle(b), U, da, db))
        col += b.ncells
        sh += ns
    end
end

function get_strengths!(body)
    for b in components(body)
        g = zeros(b.ncells)
        check(ccall((:fp_body_get_strength, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}), ctx(), handle(b), g))
        b.strength[:, 1] .= g
        pnl._solvedflag(b, true)
    end
end

"""
    solve(body, Uinfs[, Das, Dbs]; solver=SOLVER_LU, elprescribe=nothing, opts...)

Device assembly, no-flow-through right-hand side and linear solve; the
strengths are copied back into `body.strength[:, 1]`.  A watertight
`RigidWakeBody{VortexRing, 2}` goes through the least-squares formulation with
`elprescribe` (default: `pnl.calc_elprescribe(body)`).
"""
function solve(body, Uinfs::AbstractMatrix, Das=zeros(3, 0), Dbs=zeros(3, 0);
               solver::Integer=SOLVER_LU, elprescribe=nothing, opts...)
    set_freestream!(body, Uinfs, Das, Dbs)
    o = Ref(solve_opts(solver; opts...))
    s = Ref{SolveStats}()
    hs = handles(body)
    lsq = any(b -> b isa pnl.RigidWakeBody{pnl.VortexRing,2}, components(body))
    if lsq || elprescribe !== nothing
        pres = elprescribe === nothing ? pnl.calc_elprescribe(body) : elprescribe
        idx = Int64[p[1] for p in pres]
        val = Float64[p[2] for p in pres]
        check(ccall((:fp_solve_bodies_lsq, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Int64}, Ptr{Float64}, Int64, Cint, Ref{SolveOpts},
                     Ref{SolveStats}), ctx(), hs, length(hs), idx, val, length(idx), 1, o, s))
    else
        check(ccall((:fp_solve_bodies, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ref{SolveOpts}, Ref{SolveStats}),
                    ctx(), hs, length(hs), o, s))
    end
    get_strengths!(body)
    return s[]
end

"""Uind!(body, targets, out): out (3 x nt) += velocity induced by the solved body."""
function Uind!(body, targets::AbstractMatrix{Float64}, out::AbstractMatrix{Float64})
    hs = handles(body)
    check(ccall((:fp_uind, lib), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Float64}, Int64, Ptr{Float64}),
                ctx(), hs, length(hs), targets, size(targets, 2), out))
    return out
end

"""phi!(body, targets, out): out (nt) += potential of the solved body."""
function phi!(body, targets::AbstractMatrix{Float64}, out::AbstractVector{Float64})
    hs = handles(body)
    check(ccall((:fp_phi, lib), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Float64}, Int64, Ptr{Float64}),
                ctx(), hs, length(hs), targets, size(targets, 2), out))
    return out
end

"""calcfield_U!(out, sourcebody, targetbody): Uinf + induced velocity at the control points (3 x N, +=)."""
function calcfield_U!(out::AbstractMatrix{Float64}, sourcebody, targetbody=sourcebody)
    if targetbody === sourcebody
        hs = handles(sourcebody)
        check(ccall((:fp_calcfield_U, lib), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Float64}),
                    ctx(), hs, length(hs), out))
    else
        normals = pnl.calc_normals(targetbody)
        out .+= hcat(pnl.get_field(targetbody, "Uinf")["field_data"]...)
        Uind!(sourcebody, pnl.calc_controlpoints(targetbody, normals), out)
    end
    return out
end

"""calcfield_Cp!(out, body, Us, Uref): out += 1 - (|U|/Uref)^2 with the Kutta average."""
function calcfield_Cp!(out::AbstractVector{Float64}, body, Us::AbstractMatrix{Float64}, Uref::Real;
                       correct_kuttacondition=true, kutta_mode=KUTTA_REFERENCE)
    col = 0
    for b in components(body)
        lifting = !(b isa pnl.NonLiftingBody) && correct_kuttacondition
        sh = lifting ? Int64.(b.shedding) : zeros(Int64, 6, 0)
        o = zeros(b.ncells)
        check(ccall((:fp_calcfield_cp, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Int64, Cdouble, Ptr{Int64}, Int64, Cint, Cint, Ptr{Float64}),
                    ctx(), Us[:, col+1:col+b.ncells], b.ncells, Uref, sh, size(sh, 2), 1,
                    lifting ? kutta_mode : KUTTA_NONE, o))
        out[col+1:col+b.ncells] .+= o
        col += b.ncells
    end
    return out
end

const SHARPTE_OFF = Cint(0)
const SHARPTE_PAIRS = Cint(1)
const SHARPTE_CELLS = Cint(2)

"""calcfield_Ugradmu!(out, body; force_cellTE, anglecrit, sharpTE): combined -1/2 grad(mu) (3 x N, +=).
`sharpTE=true` uses the reference's structured-grid layout (the two triangles of each trailing-edge quad)."""
function calcfield_Ugradmu!(out::AbstractMatrix{Float64}, body; force_cellTE=true, anglecrit=30.0, sharpTE=false)
    hs = handles(body)
    check(ccall((:fp_calcfield_ugradmu_combined, lib), Cint,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Cint, Cdouble, Cint, Ptr{Float64}),
                ctx(), hs, length(hs), force_cellTE, anglecrit, sharpTE ? SHARPTE_PAIRS : SHARPTE_OFF, out))
    return out
end

"""calcfield_Ugradmu_node!(out, body; sharpTE): node-based -1/2 grad(mu) (3 x N, +=); returns the nodal values."""
function calcfield_Ugradmu_node!(out::AbstractMatrix{Float64}, body; sharpTE=false)
    hs = handles(body)
    nodal = zeros(sum(b.nnodes for b in components(body)))
    check(ccall((:fp_calcfield_ugradmu_node, lib), Cint,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Cint, Ptr{Float64}, Ptr{Float64}),
                ctx(), hs, length(hs), sharpTE ? SHARPTE_PAIRS : SHARPTE_OFF, out, nodal))
    return nodal
end

"""
    calcfield_F!(out, body, Cps, Uinf, rho; Xac, correct_kuttacondition) -> (Ftot, Mtot)

Element forces (3 x N; the trailing-edge cells overwritten by the Kutta override, the others added to),
with the integrated force and the moment about `Xac` (src/FLOWPanel_postprocess.jl:1199-1307, 1543-1784).
"""
function calcfield_F!(out::AbstractMatrix{Float64}, body, Cps::AbstractVector{Float64}, Uinf::Real, rho::Real;
                      Xac=zeros(3), correct_kuttacondition=true, kutta_mode=KUTTA_REFERENCE)
    hs = handles(body)
    Ftot = zeros(3)
    Mtot = zeros(3)
    check(ccall((:fp_calcfield_F, lib), Cint,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Float64}, Cdouble, Cdouble, Cint, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ctx(), hs, length(hs), Cps, Uinf, rho, correct_kuttacondition ? kutta_mode : KUTTA_NONE,
                Vector{Float64}(Xac), C_NULL, out, Ftot, Mtot))
    return Ftot, Mtot
end

"""calcfield_sectionalforce(body, Fs, nchord, nspan; spandirection): force per unit span and station positions
of a structured grid (src/FLOWPanel_postprocess.jl:1411-1542)."""
function calcfield_sectionalforce(body, Fs::AbstractMatrix{Float64}, nchord::Integer, nspan::Integer;
                                  spandirection=[0.0, 1.0, 0.0])
    outf = zeros(3, nspan)
    pos = zeros(nspan)
    check(ccall((:fp_calcfield_sectionalforce, lib), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Int64, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}), ctx(), handle(body), Fs, C_NULL, nchord, nspan, Vector{Float64}(spandirection),
                outf, pos))
    return outf, pos
end

# ---- the solver-function protocol on caller matrices ---------------------------
function solve_matrix!(y::AbstractVecOrMat{Float64}, A::AbstractMatrix{Float64}, b::AbstractVecOrMat{Float64},
                       solver::Integer; opts...)
    m = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fp_mat_upload, lib), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ref{Ptr{Cvoid}}),
                ctx(), size(A, 1), A, m))
    o = Ref(solve_opts(solver; opts...))
    s = Ref{SolveStats}()
    st = ccall((:fp_solve_mat, lib), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Ref{SolveOpts}, Ref{SolveStats}),
               ctx(), m[], b, size(b, 2), y, o, s)
    ccall((:fp_mat_destroy, lib), Cvoid, (Ptr{Cvoid},), m[])
    check(st)
    return y
end

solve_ludiv!(y, A, b; opts...) = solve_matrix!(y, A, b, SOLVER_LU; opts...)
solve_backslash!(y, A, b; opts...) = solve_matrix!(y, A, b, SOLVER_LU; opts...)
solve_gmres!(y, A, b; atol=sqrt(eps()), rtol=atol, memory=50, opts...) =
    solve_matrix!(y, A, b, SOLVER_GMRES; tol=rtol, restart=memory, opts...)
solve_lu32_ir!(y, A, b; opts...) = solve_matrix!(y, A, b, SOLVER_LU32_IR; opts...)
solve_gmres_ir!(y, A, b; opts...) = solve_matrix!(y, A, b, SOLVER_GMRES_IR; opts...)

# ---- many freestreams as right-hand sides (fixed-wake convention) --------------
function solve_freestreams(body, Uinfs3k::Array{Float64,3}, Das, Dbs; solver::Integer=SOLVER_LU32_IR, opts...)
    k = size(Uinfs3k, 3)
    set_freestream!(body, Uinfs3k[:, :, 1], Das, Dbs)
    hs = handles(body)
    o = Ref(solve_opts(solver; opts...))
    G = zeros(body.ncells, k)
    s = Ref{SolveStats}()
    check(ccall((:fp_solve_bodies_multi, lib), Cint,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Float64}, Int64, Ref{SolveOpts}, Ptr{Float64},
                 Ref{SolveStats}), ctx(), hs, length(hs), Uinfs3k, k, o, G, s))
    return G
end

# ---- batches of independent bodies (one captured hipGraph) ---------------------
mutable struct Batch
    h::Ptr{Cvoid}
    bodies::Vector{Any}
end

function batch(bodies; solver::Integer=SOLVER_LU32_IR, opts...)
    o = Ref(solve_opts(solver; opts...))
    b = Ref{Ptr{Cvoid}}(C_NULL)
    hs = Ptr{Cvoid}[handle(x) for x in bodies]
    check(ccall((:fp_batch_create, lib), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ref{SolveOpts}, Ref{Ptr{Cvoid}}),
                ctx(), hs, length(hs), o, b))
    x = Batch(b[], collect(bodies))
    finalizer(y -> ccall((:fp_batch_destroy, lib), Cvoid, (Ptr{Cvoid},), y.h), x)
    return x
end

function solve!(x::Batch)
    s = Ref{BatchStats}()
    check(ccall((:fp_batch_solve, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{BatchStats}), ctx(), x.h, s))
    foreach(get_strengths!, x.bodies)
    return s[]
end

# ---- lift/drag of the offset-probe workflow and its adjoint --------------------
function liftdrag_opts(; offset=0.02, length=1.0, S=1.0, solver::Integer=SOLVER_GMRES_IR, span=(0.0, 1.0, 0.0))
    o = Ref{LiftDragOpts}()
    ccall((:fp_liftdrag_opts_default, lib), Cvoid, (Ref{LiftDragOpts}, Cint), o, solver)
    x = o[]
    return LiftDragOpts(offset, length, x.probe_charlen, x.kutta_mode, span, S, x.solve)
end

function set_nodes!(body, nodes::Matrix{Float64})
    check(ccall((:fp_body_set_nodes, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}), ctx(), handle(body), nodes))
end

"""(CL, CDi) as a function of the node coordinates and Uinf; the connectivity is fixed."""
function lift_drag(body, nodes::Matrix{Float64}, Uinf::Vector{Float64}; opts=liftdrag_opts())
    set_nodes!(body, nodes)
    hs = handles(body)
    cl = Ref{Cdouble}()
    cd = Ref{Cdouble}()
    s = Ref{LiftDragStats}()
    check(ccall((:fp_lift_drag, lib), Cint,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Float64}, Ref{LiftDragOpts}, Ref{Cdouble}, Ref{Cdouble},
                 Ref{LiftDragStats}), ctx(), hs, length(hs), Uinf, Ref(opts), cl, cd, s))
    return (cl[], cd[])
end

function ChainRulesCore.rrule(::typeof(lift_drag), body, nodes::Matrix{Float64}, Uinf::Vector{Float64};
                              opts=liftdrag_opts())
    set_nodes!(body, nodes)
    hs = handles(body)
    gl = zeros(3, size(nodes, 2))
    gd = zeros(3, size(nodes, 2))
    ul = zeros(3)
    ud = zeros(3)
    s = Ref{LiftDragStats}()
    check(ccall((:fp_lift_drag_adjoint, lib), Cint,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Float64}, Ref{LiftDragOpts}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ref{LiftDragStats}),
                ctx(), hs, length(hs), Uinf, Ref(opts), gl, gd, ul, ud, s))
    function pullback(dy)
        dCL, dCD = dy
        return (ChainRulesCore.NoTangent(), ChainRulesCore.NoTangent(),
                dCL .* gl .+ dCD .* gd, dCL .* ul .+ dCD .* ud)
    end
    return (s[].CL, s[].CDi), pullback
end

end # module
