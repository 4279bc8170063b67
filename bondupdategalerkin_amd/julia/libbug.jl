# Julia binding of libbug.so (the C ABI declared in include/bug.h) for the reference's BUG step.
#
# Meant to be `include`d inside the reference's module after its TensorNetworks files
# (src/BUG/Core/TensorNetworks/symmetric_mps.jl, sectors.jl), so `Telum`, `SymMPS`, `MPO`,
# `_sector_block` and `cbe_bug_step!`'s keyword names resolve there.  Julia is not installed where
# this backend is built and tested, so this file has not been run; the Python package drives the
# same entry points through ctypes (bondupdategalerkin_amd/_lib.py, cstep.py) and its tests are
# the executable record of the argument meanings used below.
#
# Two levels are bound:
#   * the whole step: DeviceState / DeviceMPO handles and `cbe_bug_step_device!`, a drop-in for
#     `cbe_bug_step!(psi, mpo, tau; ...)` (bug_step.jl:17-148) whose state stays on the GPU;
#   * the operator-action seams (dd_cbe.jl:9-117, mpo_environments.jl:65-118, krylov.jl,
#     svd.jl:1813-1921) over caller-owned device buffers (raw pointers, e.g. `pointer(x)` of an
#     AMDGPU.jl `ROCArray{ComplexF64}`), for a maintainer who keeps the step's host logic in
#     Julia and moves only the contractions.

const LIBBUG = get(ENV, "BUG_LIBBUG", joinpath(@__DIR__, "..", "libbug.so"))

const BUG_OK = Cint(0)
const BUG_EINVAL = Cint(1)
const BUG_ENONFINITE = Cint(2)
const BUG_EINVARIANT = Cint(3)
const BUG_EDEVICE = Cint(4)
const BUG_ENOMEM = Cint(5)

const BUG_PREC_FP64 = Int32(0)
const BUG_PREC_FP32 = Int32(1)

const BUG_UNMET_NAMES = ((1, :sample), (2, :krylov), (4, :closure), (8, :projection),
                         (16, :expv), (32, :trunc))
const BUG_CEIL_RULES = Dict(:bond => Int32(0), :maxdim => Int32(1), :bond_strict => Int32(2),
                            :frame_root => Int32(3))
const BUG_KRYLOV_MODES = Dict(:auto => Int32(0), :lanczos => Int32(1), :arnoldi => Int32(2))
const BUG_ROOT_METHODS = Dict(:exp => Int32(0), :trapezoid => Int32(1), :radau2 => Int32(2),
                              :radau3 => Int32(3))

# ── C structs (field for field as in include/bug.h; Julia lays isbits fields out as C does) ──

struct BugLeg
    nsec::Int32
    charge::Ptr{Int32}
    dim::Ptr{Int32}
end

struct BugBsm
    row::BugLeg
    col::BugLeg
    nslice::Int32
    shift::Ptr{Int32}
    offset::Ptr{Int64}
    data::Ptr{Cvoid}
    nelem::Int64
end

struct BugMpoSite
    wl::Int32
    wr::Int32
    cl::Ptr{Int32}
    cr::Ptr{Int32}
    nent::Int32
    a::Ptr{Int32}
    b::Ptr{Int32}
    sout::Ptr{Int32}
    sin::Ptr{Int32}
    coeff::Ptr{Float64}
end

mutable struct BugStepOpts
    order::Int32
    maxdim::Int32
    hermitian::Int32
    work_ceiling_rule::Int32
    bond_ceiling_factor::Float64
    root_ceiling_factor::Float64
    tol_basis::Float64
    tol_trunc::Float64
    tol_expv::Float64
    tol_krylov::Float64
    root::Int32
    reorth::Int32
    truncate::Int32
    precision::Int32
    parallel::Int32
    seed::UInt64
    fixed_width::Int32
    tail_factor::Float64
    krylov_mode::Int32
    mode::Int32
    exact::Int32
    fold_omega::Int32
    krylov_rank_tol::Float64
    root_method::Int32
    root_shift::Int32
    root_maxiter::Int32
    BugStepOpts() = new()
end

struct BugBondLog
    bond::Int32
    side::Int32
    sample::Float64
    tail::Float64
    krylov_c::Float64
    krylov_d::Float64
    n_probes::Int32
    n_holdout::Int32
    n_selected::Int32
    n_targeted::Int32
    n_krylov::Int32
    n_h1::Int32
    workcap::Int32
    met::Int32
    unmet::Int32
    width::Int32
    r_old::Int32
    r_final::Int32
    discarded::Float64
end

mutable struct BugStepInfo
    tolerance_met::Int32
    unmet::Int32
    max_sample::Float64
    max_tail::Float64
    max_krylov::Float64
    max_krylov_c::Float64
    max_krylov_d::Float64
    discarded::Float64
    worst_discarded::Float64
    expv_err::Float64
    expv_iters::Int32
    proj_err::Float64
    defect::Float64
    leak::NTuple{4,Float64}
    n_probes::Int64
    n_selected::Int64
    n_krylov::Int64
    n_h1::Int64
    n_h1_closure::Int64
    n_closure_sweeps::Int64
    workcap_reached::Int32
    maxdim_reached::Int32
    r_old::Int32
    r_aug::Int32
    r_final::Int32
    root_link::Int32
    root_rank::Int32
    root_applications::Int32
    seconds::NTuple{6,Float64}
    expanded::Ptr{Int32}
    expanded_cap::Int32
    log::Ptr{BugBondLog}
    log_cap::Int32
    log_count::Int32
    BugStepInfo() = new()
end

struct BugExpvReport
    iterations::Int32
    substeps::Int32
    err::Float64
end

# ── context ─────────────────────────────────────────────────────────────────────────────────

mutable struct BugCtx
    h::Ptr{Cvoid}
end

function BugCtx(device::Integer = 0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:bug_ctx_create, LIBBUG), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, r)
    rc == BUG_OK || error("bug_ctx_create failed with status $rc")
    ctx = BugCtx(r[])
    finalizer(c -> ccall((:bug_ctx_destroy, LIBBUG), Cint, (Ptr{Cvoid},), c.h), ctx)
    return ctx
end

"Throw the library's message for a failed call; invalid arguments become `ArgumentError`, as the
reference throws before touching the state (bug_step_stages.jl:41-111)."
function bugcheck(ctx::BugCtx, rc::Cint)
    rc == BUG_OK && return nothing
    msg = unsafe_string(ccall((:bug_last_error, LIBBUG), Cstring, (Ptr{Cvoid},), ctx.h))
    rc == BUG_EINVAL && throw(ArgumentError(msg))
    error("libbug status $rc: $msg")
end

# ── host families from the reference's tensors ─────────────────────────────────────────────

"The integer U(1) charge (2S^z for spin-1/2) of a Telum sector label."
u1_charge(label) = Int32(label[1][1])

"""
    HostFamily

A site tensor `(link_l, site, link_r)` as one `bug_bsm` with HOST data: slices are the local
states (shift = their charge), blocks column-major, packed slice by slice with the row sectors in
charge order.  Keeps the arrays the descriptor points into alive.
"""
struct HostFamily
    rq::Vector{Int32}
    rd::Vector{Int32}
    cq::Vector{Int32}
    cd::Vector{Int32}
    shift::Vector{Int32}
    offset::Vector{Int64}
    data::Vector{ComplexF64}
end

_sorted_leg(space, charge) = sort([(charge(q), Int32(d)) for (q, d) in space]; by = first)

"""
    host_family(A; charge = u1_charge) -> HostFamily

Every sector `(cl, s, cr)` of `A` (`Telum.logical_qlabels`, values by `_sector_block`) must
satisfy `charge(cr) == charge(cl) + charge(s)`; a tensor whose arrows give another rule is
refused rather than reinterpreted.
"""
function host_family(A; charge = u1_charge)
    sp = Telum.spaces(A)
    rows = _sorted_leg(sp[1], charge)
    cols = _sorted_leg(sp[3], charge)
    rq, rd = first.(rows), last.(rows)
    cq, cd = first.(cols), last.(cols)
    shift = Int32[]
    slice0 = Dict{Int32,Int}()
    for (q, d) in sp[2]
        slice0[charge(q)] = length(shift)
        append!(shift, fill(charge(q), d))
    end
    nsec = length(rq)
    offset = fill(Int64(-1), length(shift) * nsec)
    pos = Int64(0)
    for k = 1:length(shift), i = 1:nsec
        j = findfirst(==(rq[i] + shift[k]), cq)
        j === nothing && continue
        offset[(k-1)*nsec+i] = pos
        pos += Int64(rd[i]) * cd[j]
    end
    data = zeros(ComplexF64, max(pos, 1))
    for (slot, ql) in pairs(Telum.logical_qlabels(A))
        cl, cs, cr = charge(ql[1]), charge(ql[2]), charge(ql[3])
        cr == cl + cs || throw(ArgumentError(
            "sector ($cl, $cs, $cr) breaks cr = cl + s; pass a `charge` that matches the arrows"))
        B = Array(_sector_block(A, slot))
        i = findfirst(==(cl), rq)
        for p = 1:size(B, 2)
            k = slice0[cs] + p
            off = offset[(k-1)*nsec+i]
            m = @view B[:, p, :]
            data[off+1:off+length(m)] .= vec(m)
        end
    end
    return HostFamily(rq, rd, cq, cd, shift, offset, data)
end

"The descriptor of `f` (valid while `f` is alive: wrap uses in `GC.@preserve f`)."
descriptor(f::HostFamily, data::Ptr{Cvoid} = Ptr{Cvoid}(pointer(f.data))) = BugBsm(
    BugLeg(length(f.rq), pointer(f.rq), pointer(f.rd)),
    BugLeg(length(f.cq), pointer(f.cq), pointer(f.cd)),
    length(f.shift), pointer(f.shift), pointer(f.offset), data, length(f.data))

"""
    HostMpoSite

An MPO tensor `(w_l, s_in, s_out, w_r)` (the leg order `_mpo_left_step` contracts,
mpo_environments.jl:65-81) as the sparse operator-valued matrix of `bug_mpo_site`: one entry per
non-zero element, channel charges from the two link legs.
"""
struct HostMpoSite
    cl::Vector{Int32}
    cr::Vector{Int32}
    a::Vector{Int32}
    b::Vector{Int32}
    sout::Vector{Int32}
    sin::Vector{Int32}
    coeff::Vector{Float64}
end

function host_mpo_site(W; charge = u1_charge)
    sp = Telum.spaces(W)
    leg_starts(space) = (starts = Dict{Int32,Int}(); n = 0;
                         for (q, d) in space; starts[charge(q)] = n; n += d; end; (starts, n))
    (ls, wl), (rs, wr) = leg_starts(sp[1]), leg_starts(sp[4])
    (ss, _) = leg_starts(sp[2])
    (so, _) = leg_starts(sp[3])
    cl = zeros(Int32, wl)
    cr = zeros(Int32, wr)
    for (q, d) in sp[1]; cl[ls[charge(q)]+1:ls[charge(q)]+d] .= charge(q); end
    for (q, d) in sp[4]; cr[rs[charge(q)]+1:rs[charge(q)]+d] .= charge(q); end
    a, b, sout, sin, coeff = Int32[], Int32[], Int32[], Int32[], Float64[]
    for (slot, ql) in pairs(Telum.logical_qlabels(W))
        qa, qi, qo, qb = charge.(ql)
        qb == qa + qo - qi || throw(ArgumentError(
            "MPO sector ($qa, $qi, $qo, $qb) breaks cr = cl + q(sout) - q(sin)"))
        B = Array(_sector_block(W, slot))
        for I in CartesianIndices(B)
            iszero(B[I]) && continue
            push!(a, ls[qa] + I[1] - 1)
            push!(sin, ss[qi] + I[2] - 1)
            push!(sout, so[qo] + I[3] - 1)
            push!(b, rs[qb] + I[4] - 1)
            push!(coeff, real(B[I]), imag(B[I]))
        end
    end
    return HostMpoSite(cl, cr, a, b, sout, sin, coeff)
end

descriptor(w::HostMpoSite) = BugMpoSite(
    length(w.cl), length(w.cr), pointer(w.cl), pointer(w.cr), length(w.a),
    pointer(w.a), pointer(w.b), pointer(w.sout), pointer(w.sin), pointer(w.coeff))

# ── the whole step ─────────────────────────────────────────────────────────────────────────

mutable struct DeviceMPO
    h::Ptr{Cvoid}
    L::Int
end

function DeviceMPO(mpo; charge = u1_charge)
    sites = [host_mpo_site(mpo[i]; charge) for i = 1:length(mpo)]
    descs = [descriptor(w) for w in sites]
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = GC.@preserve sites ccall((:bug_mpo_create, LIBBUG), Cint,
        (Int32, Ptr{BugMpoSite}, Ptr{Ptr{Cvoid}}), length(descs), descs, r)
    rc == BUG_OK || error("bug_mpo_create failed with status $rc")
    m = DeviceMPO(r[], length(mpo))
    finalizer(x -> ccall((:bug_mpo_destroy, LIBBUG), Cint, (Ptr{Cvoid},), x.h), m)
    return m
end

mutable struct DeviceState
    h::Ptr{Cvoid}
    ctx::BugCtx
    L::Int
end

"""
    DeviceState(ctx, psi::SymMPS; charge = u1_charge, stream = C_NULL)

`copy(psi)` on the GPU (`bug_state_create`, host data uploaded once); the centre is converted
to the library's 0-based index.
"""
function DeviceState(ctx::BugCtx, psi; charge = u1_charge, stream::Ptr{Cvoid} = C_NULL)
    L = length(psi)
    fams = [host_family(psi[i]; charge) for i = 1:L]
    local_charges = fams[1].shift
    all(f -> f.shift == local_charges, fams) ||
        throw(ArgumentError("sites carry different local spaces"))
    descs = [descriptor(f) for f in fams]
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = GC.@preserve fams local_charges ccall((:bug_state_create, LIBBUG), Cint,
        (Ptr{Cvoid}, Int32, Int32, Ptr{Int32}, Ptr{BugBsm}, Int32, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
        ctx.h, L, length(local_charges), local_charges, descs, psi.center - 1, stream, r)
    bugcheck(ctx, rc)
    st = DeviceState(r[], ctx, L)
    finalizer(x -> ccall((:bug_state_destroy, LIBBUG), Cint, (Ptr{Cvoid},), x.h), st)
    return st
end

"<S^z_j> at every site (`bug_state_expect_diagonal` with diag = (1/2, -1/2) on the local
states in the order the state was created with)."
function sz_profile(st::DeviceState; diag = [0.5, -0.5])
    out = zeros(Float64, st.L)
    bugcheck(st.ctx, ccall((:bug_state_expect_diagonal, LIBBUG), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), st.ctx.h, st.h, diag, out))
    return out
end

"Bond dimensions and centre of the device state, and one site's blocks on the host."
function download_site(st::DeviceState, site::Integer)
    rn, cn, ne = Ref{Int32}(0), Ref{Int32}(0), Ref{Int64}(0)
    ccall((:bug_state_site_shape, LIBBUG), Cint,
          (Ptr{Cvoid}, Int32, Ref{Int32}, Ref{Int32}, Ref{Int64}), st.h, site - 1, rn, cn, ne)
    rq, rd = zeros(Int32, rn[]), zeros(Int32, rn[])
    cq, cd = zeros(Int32, cn[]), zeros(Int32, cn[])
    d = Ref{Int32}(0)
    ccall((:bug_state_info, LIBBUG), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ref{Int32}, Ptr{Int32}),
          st.h, C_NULL, d, C_NULL)
    offset = zeros(Int64, d[] * rn[])
    data = zeros(ComplexF64, max(ne[], 1))
    bugcheck(st.ctx, ccall((:bug_state_download_site, LIBBUG), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
         Ptr{Int64}, Ptr{ComplexF64}), st.ctx.h, st.h, site - 1, rq, rd, cq, cd, offset, data))
    return (; row = collect(zip(rq, rd)), col = collect(zip(cq, cd)), offset, data)
end

"The options POD from `cbe_bug_step!`'s keywords (`root` 1-based, as in the reference)."
function step_options(; order::Int = 1, maxdim::Union{Nothing,Int} = nothing,
                      hermitian::Bool = true, work_ceiling_rule::Symbol = :bond,
                      bond_ceiling_factor = nothing, root_ceiling_factor::Real = 1.2,
                      tol_basis = nothing, tol_trunc = nothing, tol_expv = nothing,
                      tol_root = nothing, tol_krylov = nothing,
                      root::Union{Nothing,Int} = nothing, reorth::Bool = true,
                      truncate::Bool = true, precision::Symbol = :fp64,
                      parallel::Bool = false, rng::Union{Nothing,Integer} = nothing,
                      fixed_width::Union{Nothing,Int} = nothing,
                      tail_factor::Union{Nothing,Real} = nothing,
                      krylov_mode::Symbol = :auto, root_method::Symbol = :exp,
                      root_shift::Bool = true, root_maxiter::Union{Nothing,Int} = nothing)
    tol_expv !== nothing && tol_root !== nothing &&
        throw(ArgumentError("tol_root and tol_expv name the same tolerance; pass one"))
    haskey(BUG_CEIL_RULES, work_ceiling_rule) ||
        throw(ArgumentError("work_ceiling_rule must be one of $(keys(BUG_CEIL_RULES))"))
    precision in (:fp64, :mixed) || throw(ArgumentError("precision must be :fp64 or :mixed"))
    o = BugStepOpts()
    ccall((:bug_step_opts_default, LIBBUG), Cvoid, (Ref{BugStepOpts},), o)
    o.order = order
    maxdim === nothing || (o.maxdim = maxdim)
    o.hermitian = hermitian
    o.work_ceiling_rule = BUG_CEIL_RULES[work_ceiling_rule]
    bond_ceiling_factor === nothing || (o.bond_ceiling_factor = bond_ceiling_factor)
    o.root_ceiling_factor = root_ceiling_factor
    te = tol_root === nothing ? tol_expv : tol_root
    tol_basis === nothing || (o.tol_basis = tol_basis)
    tol_trunc === nothing || (o.tol_trunc = tol_trunc)
    te === nothing || (o.tol_expv = te)
    tol_krylov === nothing || (o.tol_krylov = tol_krylov)
    root === nothing || (o.root = root - 1)
    o.reorth = reorth
    o.truncate = truncate
    o.precision = precision === :mixed ? BUG_PREC_FP32 : BUG_PREC_FP64
    o.parallel = parallel
    rng === nothing || (o.seed = UInt64(rng))
    fixed_width === nothing || (o.fixed_width = fixed_width)
    tail_factor === nothing || (o.tail_factor = tail_factor)
    o.krylov_mode = BUG_KRYLOV_MODES[krylov_mode]
    haskey(BUG_ROOT_METHODS, root_method) ||
        throw(ArgumentError("root_method must be one of $(keys(BUG_ROOT_METHODS))"))
    o.root_method = BUG_ROOT_METHODS[root_method]
    o.root_shift = root_shift
    root_maxiter === nothing || (o.root_maxiter = root_maxiter)
    return o
end

"""
    cbe_bug_step_device!(st, mpo, tau; kwargs...) -> NamedTuple

`cbe_bug_step!(psi, mpo, tau; kwargs...)` (bug_step.jl:17-148) on a `DeviceState`: the same
keywords (the `DDQRExpansion(:rank_tail)` expansion; the adaptive controller, `exact`,
`keep_states`, `fold_omega = false` and `krylov_rank_tol > 0` are refused by the library with
`ArgumentError` before the state is touched), the ledger's verdict as `status` (`:met` or
`:budget_unmet`) with the failed tests in `unmet`, and the `CBEBugSweepInfo` fields.
"""
function cbe_bug_step_device!(st::DeviceState, mpo::DeviceMPO, tau::Number; kwargs...)
    o = step_options(; kwargs...)
    expanded = zeros(Int32, max(st.L - 1, 1))
    logs = Vector{BugBondLog}(undef, 4 * st.L)
    info = BugStepInfo()
    GC.@preserve expanded logs begin
        info.expanded = pointer(expanded)
        info.expanded_cap = st.L - 1
        info.log = pointer(logs)
        info.log_cap = length(logs)
        t = ComplexF64(tau)
        bugcheck(st.ctx, ccall((:bug_step, LIBBUG), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Ref{BugStepOpts},
             Ref{BugStepInfo}), st.ctx.h, st.h, mpo.h, real(t), imag(t), o, info))
    end
    unmet = [name for (bit, name) in BUG_UNMET_NAMES if info.unmet & bit != 0]
    return (; status = info.tolerance_met != 0 ? :met : :budget_unmet, unmet,
            discarded = info.worst_discarded, cumulative_discarded = info.discarded,
            expv_err = info.expv_err, expv_iters = info.expv_iters, proj_err = info.proj_err,
            leak = info.leak, max_sample = info.max_sample, max_krylov = info.max_krylov,
            n_probes = info.n_probes, n_selected = info.n_selected, n_krylov = info.n_krylov,
            n_h1 = info.n_h1, workcap_reached = info.workcap_reached != 0,
            maxdim_reached = info.maxdim_reached != 0, r_old = info.r_old, r_aug = info.r_aug,
            r_final = info.r_final, root_link = info.root_link + 1,
            expanded = expanded[1:st.L-1], bond_log = logs[1:info.log_count],
            seconds = info.seconds)
end

"""
    cbe_bug_step_batch!(states, mpos, taus; workers = 8, kwargs...)

Many independent `cbe_bug_step!` calls with one keyword set, advanced together through
`bug_step_batch`: each instance's stages before the root solve on `workers` host threads, the
root exponentials as one lockstep Krylov recurrence over grouped launches (e.g. a batch of Heston
parameter sets).  `taus` holds one step per state (a number applies to all).  Returns each
instance's `(status, expv_iters, root_link)`; full reports come from `cbe_bug_step_device!`.
"""
function cbe_bug_step_batch!(states::AbstractVector{DeviceState}, mpos::AbstractVector{DeviceMPO},
                             taus; workers::Int = 8, kwargs...)
    k = length(states)
    length(mpos) == k || throw(ArgumentError("one MPO per state"))
    ts = taus isa Number ? fill(ComplexF64(taus), k) : ComplexF64.(collect(taus))
    length(ts) == k || throw(ArgumentError("one tau per state"))
    k == 0 && return NamedTuple[]
    o = step_options(; kwargs...)
    tv = Vector{Float64}(undef, 2k)
    for i in 1:k
        tv[2i-1], tv[2i] = real(ts[i]), imag(ts[i])
    end
    infos = [BugStepInfo() for _ in 1:k]
    sh = [s.h for s in states]
    mh = [m.h for m in mpos]
    ctx = states[1].ctx
    bugcheck(ctx, ccall((:bug_step_batch, LIBBUG), Cint,
        (Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Float64}, Ref{BugStepOpts},
         Ptr{BugStepInfo}, Int32), ctx.h, k, sh, mh, tv, o, infos, workers))
    return [(; status = inf.tolerance_met != 0 ? :met : :budget_unmet, expv_iters = inf.expv_iters,
             root_link = inf.root_link + 1) for inf in infos]
end

# ── the operator-action seams over caller-owned device buffers ─────────────────────────────

"`apply_one_site` (effective_operators.jl:39-64,118-129): Y = (H_eff - shift) X."
apply_one_site!(ctx::BugCtx, L::BugBsm, W::BugMpoSite, R::BugBsm, X::BugBsm, Y::BugBsm;
                shift = 0.0, precision = BUG_PREC_FP64, stream = C_NULL) =
    bugcheck(ctx, ccall((:bug_apply_one_site, LIBBUG), Cint,
        (Ptr{Cvoid}, Ref{BugBsm}, Ref{BugMpoSite}, Ref{BugBsm}, Ref{BugBsm}, Ref{BugBsm},
         Cdouble, Int32, Ptr{Cvoid}), ctx.h, L, W, R, X, Y, shift, precision, stream))

"`apply_zero_site` (effective_operators.jl:103-107): Y = (H_0 - shift) C."
apply_zero_site!(ctx::BugCtx, L::BugBsm, R::BugBsm, C::BugBsm, Y::BugBsm;
                 shift = 0.0, precision = BUG_PREC_FP64, stream = C_NULL) =
    bugcheck(ctx, ccall((:bug_apply_zero_site, LIBBUG), Cint,
        (Ptr{Cvoid}, Ref{BugBsm}, Ref{BugBsm}, Ref{BugBsm}, Ref{BugBsm}, Cdouble, Int32,
         Ptr{Cvoid}), ctx.h, L, R, C, Y, shift, precision, stream))

"K root generators at once (`bug_apply_zero_site_batch`): Y[i] = (H_0,i - shifts[i]) C[i]."
apply_zero_site_batch!(ctx::BugCtx, L::Vector{BugBsm}, R::Vector{BugBsm}, C::Vector{BugBsm},
                       Y::Vector{BugBsm}; shifts = nothing, precision = BUG_PREC_FP64,
                       stream = C_NULL) =
    bugcheck(ctx, ccall((:bug_apply_zero_site_batch, LIBBUG), Cint,
        (Ptr{Cvoid}, Int32, Ptr{BugBsm}, Ptr{BugBsm}, Ptr{BugBsm}, Ptr{BugBsm}, Ptr{Float64},
         Int32, Ptr{Cvoid}), ctx.h, length(L), L, R, C, Y,
        shifts === nothing ? Ptr{Float64}(C_NULL) : Vector{Float64}(shifts), precision, stream))

"`_mpo_left_step` (mpo_environments.jl:65-81): the environment on the next link."
env_push_left!(ctx::BugCtx, L::BugBsm, A::BugBsm, W::BugMpoSite, Lout::BugBsm;
               precision = BUG_PREC_FP64, stream = C_NULL) =
    bugcheck(ctx, ccall((:bug_env_push_left, LIBBUG), Cint,
        (Ptr{Cvoid}, Ref{BugBsm}, Ref{BugBsm}, Ptr{BugBsm}, Ref{BugMpoSite}, Ref{BugBsm},
         Int32, Ptr{Cvoid}), ctx.h, L, A, C_NULL, W, Lout, precision, stream))

"`_mpo_right_step` (mpo_environments.jl:92-105)."
env_push_right!(ctx::BugCtx, R::BugBsm, B::BugBsm, W::BugMpoSite, Rout::BugBsm;
                precision = BUG_PREC_FP64, stream = C_NULL) =
    bugcheck(ctx, ccall((:bug_env_push_right, LIBBUG), Cint,
        (Ptr{Cvoid}, Ref{BugBsm}, Ref{BugBsm}, Ptr{BugBsm}, Ref{BugMpoSite}, Ref{BugBsm},
         Int32, Ptr{Cvoid}), ctx.h, R, B, C_NULL, W, Rout, precision, stream))

"""
    two_site_sketch!(ctx, side, L, W1, W2, R, A, B, Om, out; precision, stream)

`apply_two_site_sketch!` (dd_cbe.jl:70-89) through `_stream_left` / `_stream_right`
(mpo_environments.jl:409-481): `side = :left` gives (H Theta) Om' with Om in B's layout,
`:right` Om' (H Theta) with Om in A's layout, Theta = A B never formed.
"""
function two_site_sketch!(ctx::BugCtx, side::Symbol, L::BugBsm, W1::BugMpoSite, W2::BugMpoSite,
                          R::BugBsm, A::BugBsm, B::BugBsm, Om::BugBsm, out::BugBsm;
                          precision = BUG_PREC_FP64, stream = C_NULL)
    fn = side === :left ? :bug_two_site_sketch_left :
         side === :right ? :bug_two_site_sketch_right :
         throw(ArgumentError("side must be :left or :right"))
    rc = ccall((fn, LIBBUG), Cint,
        (Ptr{Cvoid}, Ref{BugBsm}, Ref{BugMpoSite}, Ref{BugMpoSite}, Ref{BugBsm}, Ref{BugBsm},
         Ref{BugBsm}, Ref{BugBsm}, Ref{BugBsm}, Int32, Ptr{Cvoid}),
        ctx.h, L, W1, W2, R, A, B, Om, out, precision, stream)
    bugcheck(ctx, rc)
end

# The action of `expv` is a Julia closure reached through a C function pointer; the closure
# travels as the `user` pointer.
function _apply_trampoline(user::Ptr{Cvoid}, x::Ptr{Cvoid}, y::Ptr{Cvoid}, stream::Ptr{Cvoid})
    f = unsafe_pointer_to_objref(user)
    try
        f(x, y, stream)
        return Cint(0)
    catch
        return Cint(1)
    end
end

"""
    expv!(ctx, apply, n, tau, v, out; hermitian, tol, maxiter, reorth, conv_tol, stream)

`expv` (root_methods.jl:359-402; `lanczos_expv`/`arnoldi_expv`, krylov.jl:128-263,343-442):
out = exp(tau A) v for device vectors of n ComplexF64, with `apply(x, y, stream)` enqueueing
y = A x (for the root solve: `apply_zero_site!` or `apply_one_site!` on descriptors pointing at
x and y).  Returns (iterations, substeps, err).
"""
function expv!(ctx::BugCtx, apply, n::Integer, tau::Number, v::Ptr{Cvoid}, out::Ptr{Cvoid};
               hermitian::Bool = true, tol::Real = hermitian ? 1e-15 : 1e-13,
               maxiter::Integer = 30, reorth::Bool = true, conv_tol::Real = 0.0,
               stream = C_NULL)
    cb = @cfunction(_apply_trampoline, Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}))
    rep = Ref(BugExpvReport(0, 0, 0.0))
    t = ComplexF64(tau)
    GC.@preserve apply begin
        rc = ccall((:bug_expv, LIBBUG), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Any, Int64, Cdouble, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}, Int32,
             Cdouble, Int32, Int32, Cdouble, Ref{BugExpvReport}, Ptr{Cvoid}),
            ctx.h, cb, apply, n, real(t), imag(t), v, out, hermitian, tol, maxiter, reorth,
            conv_tol, rep, stream)
    end
    bugcheck(ctx, rc)
    return (; iterations = rep[].iterations, substeps = rep[].substeps, err = rep[].err)
end

"""
    svd_select(spectra; cutoff = 0.0, nkeep = nothing)

Telum's global selection (svd.jl:1813-1921) on per-sector spectra, each sorted descending: the
kept count per sector, every value's (sector, index) in descending order (1-based), how many of
them are kept, and the kept and truncated weights.
"""
function svd_select(spectra::AbstractVector{<:AbstractVector{<:Real}}; cutoff::Real = 0.0,
                    nkeep::Union{Nothing,Integer} = nothing)
    lens = Int32[length(s) for s in spectra]
    flat = Float64[x for s in spectra for x in s]
    keep = zeros(Int32, max(length(spectra), 1))
    order = zeros(Int32, max(2 * length(flat), 1))
    nk = Ref{Int32}(0)
    w = zeros(2)
    rc = ccall((:bug_svd_select, LIBBUG), Cint,
        (Int32, Ptr{Int32}, Ptr{Float64}, Cdouble, Int32, Ptr{Int32}, Ptr{Int32}, Ref{Int32},
         Ptr{Float64}), length(spectra), lens, flat, cutoff, nkeep === nothing ? -1 : nkeep,
        keep, order, nk, w)
    rc == BUG_OK || throw(ArgumentError("bug_svd_select refused its arguments ($rc)"))
    pairs_ = [(order[2e-1] + 1, order[2e] + 1) for e = 1:length(flat)]
    return keep[1:length(spectra)], pairs_, Int(nk[]), (w[1], w[2])
end
