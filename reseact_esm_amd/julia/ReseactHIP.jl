# ReseactHIP.jl -- Julia host side of the MI355X split-step library.
#
# Mirrors the operator API of tools/reactant_handoff/rx_traced_integrator.jl
# (PICtrl, ssprk43_step, ros23_step, adaptive_solve, *_step_vjp) and
# tools/reactant_handoff/op_split.jl (lie_trotter_solve, lie_trotter_solve_bisect)
# on top of the C ABI of
# include/reseact_hip.h, called with `ccall`.  The RHS closures of the
# reference are replaced by a `Context` that holds the grid, wind, mechanism
# and emissions on the GPU.  `rrule(split_loss, ...)` gives ChainRulesCore the
# discrete adjoint of the whole macro-step loop (checkpoint per split step,
# fixed-sequence replay, Rosenbrock stages recomputed).
#
# Not exercised in this repository's CI (no Julia toolchain in the build
# image); every call here is a one-to-one wrapper of a function whose
# behaviour the Python tests pin through the same ABI.
module ReseactHIP

using ChainRulesCore

const LIB = get(ENV, "RESEACT_HIP_LIB", joinpath(@__DIR__, "..", "libreseact_hip.so"))

# ---- controller (rx_traced_integrator.jl:86-99) -----------------------------
Base.@kwdef struct PICtrl
    beta1::Float64
    beta2::Float64
    qmin::Float64 = 1 / 5
    qmax::Float64 = 10.0
    gamma::Float64 = 9 / 10
    qsteady_min::Float64 = 1.0
    qsteady_max::Float64
    qoldinit::Float64 = 1e-4
end
pictrl_ros23() = PICtrl(; beta1 = 7 / 20, beta2 = 1 / 5, qsteady_max = 6 / 5)
pictrl_ssprk43() = PICtrl(; beta1 = 7 / 30, beta2 = 2 / 15, qsteady_max = 1.0)

struct Counts
    naccept::Int64
    nreject::Int64
    t_end::Float64
    dt_end::Float64
end
Counts() = Counts(0, 0, 0.0, 0.0)

struct SplitOpts
    ctrl_T::PICtrl
    ctrl_C::PICtrl
    atol_T::Float64
    rtol_T::Float64
    atol_C::Float64
    rtol_C::Float64
    clamp_nonneg::Cint
    chem_percell::Cint
    maxiters::Int64
end

struct AdjointStats
    naccept_T::Int64; naccept_C::Int64; nreject_T::Int64; nreject_C::Int64
    nvjp_T::Int64; nvjp_C::Int64
    seconds_forward::Float64; seconds_backward::Float64; seconds_replay::Float64
    replay_maxrel::Float64
end
AdjointStats() = AdjointStats(0, 0, 0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0)

# ---- context ----------------------------------------------------------------
mutable struct Context
    ptr::Ptr{Cvoid}
    nlon::Int; nlat::Int; nlev::Int; ns::Int; nt::Int
    function Context(; device::Integer = 0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:rs_create, LIB), Cint, (Ref{Ptr{Cvoid}}, Cint), r, device)
        rc == 0 || error("rs_create failed ($rc)")
        c = new(r[], 0, 0, 0, 0, 0)
        finalizer(x -> ccall((:rs_destroy, LIB), Cint, (Ptr{Cvoid},), x.ptr), c)
        return c
    end
end
ncells(c::Context) = c.nlon * c.nlat * c.nlev

function check(c::Context, rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:rs_last_error, LIB), Cstring, (Ptr{Cvoid},), c.ptr))
    rc == 4 && error("maxiters: $msg")              # host_adaptive! raises the same
    error("reseact_hip error $rc: $msg")
end

function set_grid!(c::Context, nlon, nlat, nlev, ns, nt)
    check(c, ccall((:rs_set_grid, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint),
                   c.ptr, nlon, nlat, nlev, ns, nt))
    c.nlon, c.nlat, c.nlev, c.ns, c.nt = nlon, nlat, nlev, ns, nt
    return c
end
set_wind!(c::Context, cx::Vector{Float64}, wz::Vector{Float64}, wl::Vector{Float64}) =
    check(c, ccall((:rs_set_wind, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   c.ptr, cx, wz, wl))
# reactants/products: 2 x NR Int32 (column r = reaction r, -1 = none), 0-based species
set_mechanism!(c::Context, reactants::Matrix{Int32}, products::Matrix{Int32},
               k::Vector{Float64}, photo::Vector{Int32}) =
    check(c, ccall((:rs_set_mechanism, LIB), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Ptr{Int32}),
                   c.ptr, length(k), reactants, products, k, photo))
set_photolysis!(c::Context, jfac::Union{Nothing,Vector{Float64}}) =
    check(c, ccall((:rs_set_photolysis, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr,
                   jfac === nothing ? C_NULL : jfac))
set_emissions!(c::Context, e::Union{Nothing,Vector{Float64}}) =
    check(c, ccall((:rs_set_emissions, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr,
                   e === nothing ? C_NULL : e))
# The reference's own transport: monotone PPM in CWC mixing-ratio form
# (species 0 of the transported block is the air mass).  Arrays are flat,
# longitude fastest, as in include/reseact_hip.h.
const TRANSPORT_UPWIND3 = Cint(0)
const TRANSPORT_PPM = Cint(1)
set_transport_scheme!(c::Context, scheme::Integer) =
    check(c, ccall((:rs_set_transport_scheme, LIB), Cint, (Ptr{Cvoid}, Cint), c.ptr, scheme))
set_ppm_geometry!(c::Context, lon_periodic::Bool, dlon::Float64, coslat_e::Vector{Float64},
                  dS_lat::Vector{Float64}, dphi_lat::Vector{Float64}) =
    check(c, ccall((:rs_set_ppm_geometry, LIB), Cint,
                   (Ptr{Cvoid}, Cint, Cdouble, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   c.ptr, lon_periodic ? 1 : 0, dlon, coslat_e, dS_lat, dphi_lat))
set_mass_fluxes!(c::Context, Mx::Vector{Float64}, My::Vector{Float64}, Mz::Vector{Float64},
                 dp::Vector{Float64}) =
    check(c, ccall((:rs_set_mass_fluxes, LIB), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   c.ptr, Mx, My, Mz, dp))
_ornull(a) = a === nothing ? C_NULL : a
set_inflow!(c::Context, qw, qe, qs, qn) =
    check(c, ccall((:rs_set_inflow, LIB), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   c.ptr, _ornull(qw), _ornull(qe), _ornull(qs), _ornull(qn)))
set_pressure_fixer!(c::Context, dPSdt::Union{Nothing,Vector{Float64}}) =
    check(c, ccall((:rs_set_pressure_fixer, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr,
                   _ornull(dPSdt)))
# Mz from continuity with the column pressure fixer (reseact.esm:262-291);
# returns (Mz, largest |scan value at the lid|) and installs Mz in the context
function diagnose_mz!(c::Context, nlon::Int, nlat::Int, nlev::Int;
                      dB::Union{Nothing,Vector{Float64}}=nothing,
                      dPSdt::Union{Nothing,Vector{Float64}}=nothing)
    Mz = Vector{Float64}(undef, (nlev + 1) * nlat * nlon)
    resid = Ref{Float64}(0.0)
    check(c, ccall((:rs_diagnose_mz, LIB), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
                   c.ptr, _ornull(dB), _ornull(dPSdt), Mz, resid))
    return Mz, resid[]
end
# pbl_f: layer fraction inside the PBL per cell (nothing: off); tau_pblmix in s
set_pbl_mixing!(c::Context, pbl_f::Union{Nothing,Vector{Float64}}, tau_pblmix::Float64=900.0) =
    check(c, ccall((:rs_set_pbl_mixing, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cdouble), c.ptr,
                   _ornull(pbl_f), tau_pblmix))
set_state!(c::Context, u::Vector{Float64}) =
    check(c, ccall((:rs_set_state, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, u))
function get_state(c::Context)
    u = Vector{Float64}(undef, c.ns * ncells(c))
    check(c, ccall((:rs_get_state, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, u))
    return u
end
function candidate(c::Context)
    u = Vector{Float64}(undef, c.ns * ncells(c))
    check(c, ccall((:rs_get_candidate, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, u))
    return u
end

# ---- step plug-ins (rx_traced_integrator.jl:279-437) --------------------------
# `f` of the reference is the context; the layout is the same species-major u.
function ssprk43_step(c::Context, u, t, dt, abstol::Float64, reltol::Float64)
    set_state!(c, u)
    e = Ref(0.0)
    check(c, ccall((:rs_ssprk43_step, LIB), Cint,
                   (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Ref{Cdouble}),
                   c.ptr, t, dt, abstol, reltol, e))
    return candidate(c), e[]
end

function ros23_step(c::Context, u, t, dt, NS::Int, NC::Int, masks, abstol::Float64,
                    reltol::Float64; kwargs...)
    (NS == c.ns && NC == ncells(c)) || error("ros23_step: layout does not match the context")
    set_state!(c, u)
    e = Ref(0.0)
    check(c, ccall((:rs_ros23_step, LIB), Cint,
                   (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Ref{Cdouble}),
                   c.ptr, t, dt, abstol, reltol, e))
    return candidate(c), e[]
end

# adaptive_solve (:455-507): `stepfn` selects the operator (ssprk43_step or ros23_step)
function adaptive_solve(stepfn, c::Context, u0, t0, tend, dt0, ctrl::PICtrl;
                        abstol = 1e-6, reltol = 1e-3, maxiters::Float64 = 2.0e4,
                        clamp_nonneg::Bool = false)
    set_state!(c, u0)
    dt = Ref(Float64(dt0)); cnt = Ref(Counts())
    sym = stepfn === ssprk43_step ? :rs_adaptive_transport :
          stepfn === ros23_step ? :rs_adaptive_chem : error("unknown step function")
    rc = ccall((sym, LIB), Cint,
               (Ptr{Cvoid}, Cdouble, Cdouble, Ref{Cdouble}, Ref{PICtrl}, Cdouble, Cdouble, Cint,
                Int64, Ref{Counts}),
               c.ptr, t0, tend, dt, Ref(ctrl), abstol, reltol, clamp_nonneg, Int64(maxiters), cnt)
    check(c, rc)
    return get_state(c), cnt[].t_end, dt[], Float64(cnt[].naccept), Float64(cnt[].nreject)
end

# ---- Lie-Trotter driver (op_split.jl:100-164, :219-242) ---------------------
const RS_CLAMP_NONNEG = Cint(1)
const RS_POSITIVE_DOMAIN = Cint(2)
const RS_E_NONFINITE = Cint(3)
const RS_E_MAXITERS = Cint(4)

clamp_state!(c::Context) = check(c, ccall((:rs_clamp_state, LIB), Cint, (Ptr{Cvoid},), c.ptr))

# Stop grid (op_split.jl:122-144): macro lattice short of tf, forcing boundaries
# in the half-open t0 < x <= tf, and tf.
function forcing_stops(t0::Float64, tf::Float64, macro_dt::Real, forcing_tstops)
    fstops = Set(round(t; digits = 6) for t in forcing_tstops if t0 + 1e-6 < t <= tf + 1e-6)
    grid = collect((t0 + macro_dt):macro_dt:(tf - 1e-9))
    return sort!(unique!(vcat(grid, collect(fstops), Float64[tf]))), fstops
end

function lie_trotter_solve(c::Context, u0, tspan; macro_dt::Real,
                           reltols = (1e-3, 1e-3), abstols = (1e-6, 1e-6),
                           refresh = (_t -> nothing), forcing_tstops = Float64[],
                           clamp_nonneg::Bool = true, positive_domain::Bool = false,
                           dt0 = (15.0, 0.5), maxiters = 20000)
    t0, tf = Float64(tspan[1]), Float64(tspan[2])
    stops, fstops = forcing_stops(t0, tf, macro_dt, forcing_tstops)
    flags = positive_domain ? RS_POSITIVE_DOMAIN : Cint(0)
    opts = SplitOpts(pictrl_ssprk43(), pictrl_ros23(), abstols[1], reltols[1], abstols[2],
                     reltols[2], flags, 0, maxiters)
    set_state!(c, u0)
    dtT = Ref(Float64(dt0[1])); dtC = Ref(Float64(dt0[2]))
    naT = nrT = naC = nrC = 0; nmacro = 0; tcur = t0; retcode = :Success
    for tnext in stops
        tnext <= tcur + 1e-9 && continue
        cT = Ref(Counts()); cC = Ref(Counts())
        rc = ccall((:rs_macro_step, LIB), Cint,
                   (Ptr{Cvoid}, Cdouble, Cdouble, Ref{Cdouble}, Ref{Cdouble}, Ref{SplitOpts},
                    Ref{Counts}, Ref{Counts}),
                   c.ptr, tcur, tnext, dtT, dtC, Ref(opts), cT, cC)
        if rc == RS_E_MAXITERS || rc == RS_E_NONFINITE
            retcode = rc == RS_E_MAXITERS ? :MaxIters : :Unstable
            break
        end
        check(c, rc)
        naT += cT[].naccept; nrT += cT[].nreject; naC += cC[].naccept; nrC += cC[].nreject
        clamp_nonneg && clamp_state!(c)
        tcur = tnext; nmacro += 1
        round(tcur; digits = 6) in fstops && refresh(tcur)
    end
    return (; u = get_state(c), retcode, naT, nrT, naC, nrC, nmacro)
end

_ltg_ok(r) = r.retcode === :Success || r.retcode === :Default

function lie_trotter_solve_bisect(c::Context, u0, tspan; macro_dt, min_macro_dt::Real = 9.0,
                                  on_bisect = (_t0, _tf) -> nothing, kwargs...)
    t0 = Float64(tspan[1]); tf = Float64(tspan[2])
    r = lie_trotter_solve(c, u0, (t0, tf); macro_dt = macro_dt, kwargs...)
    (_ltg_ok(r) || (tf - t0) <= min_macro_dt + 1e-9) && return r
    tm = 0.5 * (t0 + tf)
    on_bisect(t0, tf)
    r1 = lie_trotter_solve_bisect(c, u0, (t0, tm); macro_dt = min(Float64(macro_dt), tm - t0),
                                  min_macro_dt, on_bisect, kwargs...)
    _ltg_ok(r1) || return r1
    r2 = lie_trotter_solve_bisect(c, r1.u, (tm, tf); macro_dt = min(Float64(macro_dt), tf - tm),
                                  min_macro_dt, on_bisect, kwargs...)
    _ltg_ok(r2) || return r2
    return (; u = r2.u, retcode = r2.retcode, naT = r1.naT + r2.naT, nrT = r1.nrT + r2.nrT,
            naC = r1.naC + r2.naC, nrC = r1.nrC + r2.nrC, nmacro = r1.nmacro + r2.nmacro)
end

# ---- one-step adjoints (:610-623) ---------------------------------------------
# PPM: keep the stage states of rs_ssprk43_step for the VJP of the same step
keep_transport_stages!(c::Context, keep::Bool) =
    check(c, ccall((:rs_keep_transport_stages, LIB), Cint, (Ptr{Cvoid}, Cint), c.ptr, keep ? 1 : 0))

# A pullback of one SSPRK43 step: the forward keeps its stage states, the
# pullback (called before anything else touches the context) reuses them.
function ssprk43_step_with_pullback(c::Context, u, t, dt, abstol, reltol)
    keep_transport_stages!(c, true)
    set_state!(c, u)
    eest = Ref{Float64}(0.0)
    check(c, ccall((:rs_ssprk43_step, LIB), Cint,
                   (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Ref{Float64}),
                   c.ptr, t, dt, abstol, reltol, eest))
    unew = similar(u)
    check(c, ccall((:rs_get_candidate, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, unew))
    function pullback(lambda)
        check(c, ccall((:rs_set_lambda, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, lambda))
        check(c, ccall((:rs_ssprk43_vjp, LIB), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), c.ptr, t, dt))
        lam = similar(lambda)
        check(c, ccall((:rs_get_lambda, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, lam))
        return lam
    end
    return (unew, eest[]), pullback
end

function ssprk43_step_vjp(c::Context, u, t, dt, lambda)
    set_state!(c, u)
    check(c, ccall((:rs_set_lambda, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, lambda))
    check(c, ccall((:rs_ssprk43_vjp, LIB), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), c.ptr, t, dt))
    lam = similar(lambda)
    check(c, ccall((:rs_get_lambda, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, lam))
    return lam, nothing
end

function ros23_step_vjp(c::Context, u, t, dt, lambda)
    set_state!(c, u)
    check(c, ccall((:rs_set_lambda, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, lambda))
    check(c, ccall((:rs_zero_grad, LIB), Cint, (Ptr{Cvoid},), c.ptr))
    check(c, ccall((:rs_ros23_vjp, LIB), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), c.ptr, t, dt))
    lam = similar(lambda)
    gE = Vector{Float64}(undef, c.ns * c.nlon * c.nlat)
    check(c, ccall((:rs_get_lambda, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, lam))
    check(c, ccall((:rs_get_grad_emis, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), c.ptr, gE))
    return lam, gE
end

# ---- the split loop as a differentiable function -----------------------------
# J = sum over surface cells of u[species](t_end)^2 after the macro steps at
# `tstops`, as a function of the initial state u0 and the surface emissions E.
function split_loss(c::Context, u0::Vector{Float64}, E::Vector{Float64}, tstops::Vector{Float64},
                    opts::SplitOpts; species::Int = 0, dt0 = (15.0, 0.5))
    set_emissions!(c, E)
    set_state!(c, u0)
    dtT, dtC = Float64(dt0[1]), Float64(dt0[2])
    for m in 1:(length(tstops) - 1)
        a = Ref(dtT); b = Ref(dtC)
        check(c, ccall((:rs_macro_step, LIB), Cint,
                       (Ptr{Cvoid}, Cdouble, Cdouble, Ref{Cdouble}, Ref{Cdouble}, Ref{SplitOpts},
                        Ref{Counts}, Ref{Counts}),
                       c.ptr, tstops[m], tstops[m + 1], a, b, Ref(opts), Ref(Counts()), Ref(Counts())))
        dtT, dtC = a[], b[]
    end
    u = get_state(c)
    ncol = c.nlon * c.nlat
    off = species * ncells(c)
    return sum(abs2, view(u, (off + 1):(off + ncol)))
end

function ChainRulesCore.rrule(::typeof(split_loss), c::Context, u0::Vector{Float64},
                              E::Vector{Float64}, tstops::Vector{Float64}, opts::SplitOpts;
                              species::Int = 0, dt0 = (15.0, 0.5))
    set_emissions!(c, E)
    set_state!(c, u0)
    J = Ref(0.0)
    gu = similar(u0)
    gE = similar(E)
    st = Ref(AdjointStats())
    check(c, ccall((:rs_adjoint_run, LIB), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Float64}, Cdouble, Cdouble, Ref{SplitOpts}, Cint, Cint,
                    Ref{Cdouble}, Ptr{Float64}, Ptr{Float64}, Ref{AdjointStats}),
                   c.ptr, length(tstops) - 1, tstops, dt0[1], dt0[2], Ref(opts), 0, species, J,
                   gu, gE, st))
    pullback(Jbar) = (NoTangent(), NoTangent(), Jbar .* gu, Jbar .* gE, NoTangent(), NoTangent())
    return J[], pullback
end

end # module
