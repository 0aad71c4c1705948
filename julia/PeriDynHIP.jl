"""
    PeriDynHIP

Binds PeriDyn's explicit-dynamics hot path to libperidyn_hip.so (MI355X, gfx950)
through `ccall`.  After `set_device!()` (PeriDyn.DEVICE[] = :hip) it adds
methods at the reference's own dispatch seams:

  * force_density_T!(f, y, limits, mat, ::Type{Val{:hip}})   (general_material.jl:274-296);
    bond state is copied back into mat.general.intact after every call, so
    cal_damage and printdata (peri.jl:587-594, environment.jl:308-320) read it
  * short_range_contact!(y, f, type, bid, vol, den, RM) for
    NonLinearSpringContactModel12 / 11 (contacts.jl:189-247, 357-380); these
    replace the 7-argument method, which sets device = :cpu itself (contacts.jl:191)
  * update_contact_neighs!(y, type, RM; max_part) for the same two types
    (contacts.jl:434-476, 539-591); RM.neighs is filled from the device lists
  * run_resident!(env, N; neigh_update_freq, average_prop_freq) -- the resident,
    fused run loop (run!'s physics cadence, solvers.jl:97-187, with its printdata
    rows but without file output)
  * minimize_resident!(env, solver::QSDrag)                 (minimize.jl:105-231)
  * checkpoint(id) / restore!(id, bytes) -- resident restart state (io.jl:170-187)

The contact seams act on a device env holding the whole Env: call
`register!(env)` once after building it.  The material seam uses that env too
when the block is registered, else a device env of the block alone.

Julia is not installed in the build container, so this file is written against
the C ABI in include/peridyn.h and the Python binding (peridyn_amd/_lib.py,
which exercises the same entry points in tests) but has not been executed.
"""
module PeriDynHIP

using PeriDyn

const LIB = joinpath(@__DIR__, "..", "peridyn_amd", "libperidyn_hip.so")

const PD_BB, PD_OSB, PD_EP = Int32(0), Int32(1), Int32(2)
const PD_PREC_EXACT, PD_PREC_FP64, PD_PREC_FP32C = Int32(0), Int32(1), Int32(2)
const PD_BC_FIX, PD_BC_TOFRO = Int32(0), Int32(1)
const PD_BC_ONLYATSTART = 1

struct BlockDesc
    kind::Int32; ntype::Int32; n::Int64; type_start::Int64
    x::Ptr{Float64}; y::Ptr{Float64}; v::Ptr{Float64}; volume::Ptr{Float64}; type::Ptr{Int64}
    horizon::Float64; particle_size::Float64; max_neigh::Int64
    density::Ptr{Float64}; critical_stretch::Ptr{Float64}; bond_stiffness::Ptr{Float64}
    bulk_modulus::Ptr{Float64}; shear_modulus::Ptr{Float64}; psi::Ptr{Float64}
    deformed::Int32; reserved::Int32
end

struct ContactDesc
    types1::Ptr{Int64}; ntypes1::Int64; types2::Ptr{Int64}; ntypes2::Int64
    exponent::Float64; exponent_is_int::Int32; reserved::Int32
    stiffness::Float64; equi_dist::Float64; distance::Float64; max_neighs::Int64
end

struct BcDesc
    kind::Int32; options::Int32; mask::Ptr{UInt8}; rate::NTuple{3,Float64}
    freq::Float64; applyafter::Int64
end

mutable struct Ctx
    h::Ptr{Cvoid}
end

const CTX = Ref{Ctx}()

check(rc, ctx) = rc == 0 || error(unsafe_string(ccall((:pd_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx.h)))

function context(device::Integer=0)
    if !isassigned(CTX)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:pd_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h)
        rc == 0 || error("pd_ctx_create failed ($rc)")
        CTX[] = Ctx(h[])
    end
    CTX[]
end

"Set PeriDyn's device tag to :hip (get_valid_device would coerce it to :cpu)."
function set_device!()
    PeriDyn.DEVICE[] = :hip
    context()
end

kind_of(m::PeriDyn.BondBasedMaterial) = PD_BB
kind_of(m::PeriDyn.OrdinaryStateBasedMaterial) = PD_OSB
kind_of(m::PeriDyn.ElastoPlasticSolidMaterial) = PD_EP

# max_neigh of a block for the device family build: the reference's family has
# maxdeg + 1 rows after GeneralMaterial's truncation (general_material.jl:88-90),
# so the device build, which raises PD_E_OVERFLOW only when a degree reaches
# max_neigh, rebuilds exactly the reference's sets
family_rows(gen) = size(gen.family, 1)

# env id per material block, built on first use (family rebuilt on the GPU; the
# sets equal cal_family's, the layout is exported back on demand)
const ENVS = IdDict{Any,Int64}()

function env_for(mat, precision=PD_PREC_EXACT)
    get!(ENVS, mat) do
        ctx = context()
        gen, spc = mat.general, mat.specific
        id = Ref{Int64}(0)
        check(ccall((:pd_env_create, LIB), Cint, (Ptr{Cvoid}, Float64, Int32, Ref{Int64}),
                    ctx.h, 1.0, precision, id), ctx)
        nt = length(mat.type)
        x = Array{Float64}(gen.x); y = Array{Float64}(gen.y); v = Array{Float64}(gen.velocity)
        vol = Array{Float64}(gen.volume); typ = Array{Int64}(gen.type)
        dens = Array{Float64}(spc.density); s0 = Array{Float64}(spc.critical_stretch)
        c = hasproperty(spc, :bond_stiffness) ? Array{Float64}(spc.bond_stiffness) : zeros(nt, nt)
        K = hasproperty(spc, :shear_modulus) ? Array{Float64}(spc.bulk_modulus) : zeros(nt, nt)
        G = hasproperty(spc, :shear_modulus) ? Array{Float64}(spc.shear_modulus) : zeros(nt, nt)
        psi = hasproperty(spc, :ψ) ? Array{Float64}(spc.ψ) : zeros(nt, nt)
        GC.@preserve x y v vol typ dens s0 c K G psi begin
            d = BlockDesc(kind_of(mat), Int32(nt), size(x, 2), first(mat.type),
                          pointer(x), pointer(y), pointer(v), pointer(vol), pointer(typ),
                          gen.horizon, gen.particle_size, family_rows(gen),
                          pointer(dens), pointer(s0), pointer(c), pointer(K), pointer(G), pointer(psi),
                          Int32(all(gen.deformed)), Int32(0))
            check(ccall((:pd_env_add_block, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{BlockDesc}, Ptr{Int64}),
                        ctx.h, id[], d, C_NULL), ctx)
            check(ccall((:pd_env_commit, LIB), Cint, (Ptr{Cvoid}, Int64), ctx.h, id[]), ctx)
        end
        id[]
    end
end

# ---- seam registration: one device env per PeriDyn Env ----------------------
const SEAM_BLOCK = IdDict{Any,Tuple{Int64,Int64}}()    # material block -> (env id, block index)
const SEAM_CONTACT = IdDict{Any,Tuple{Int64,Int64}}()  # contact model -> (env id, model index)

"""
    register!(env; precision=PD_PREC_EXACT) -> Int64

Uploads `env` (blocks, contact models, BCs; `upload_env`) and routes the seam
methods of its material blocks and contact models to that device env: block k
of env.material_blocks is device block k-1, contact model k of
env.short_range_contact is device model k-1 (update_contact_acc! and
update_neighs! loop over them in this order, solvers.jl:257-269,351-361).
"""
function register!(env; precision=PD_PREC_EXACT)
    id = upload_env(env; precision=precision)
    for (k, mat) in enumerate(env.material_blocks)
        SEAM_BLOCK[mat] = (id, Int64(k - 1))
    end
    for (k, RM) in enumerate(env.short_range_contact)
        SEAM_CONTACT[RM] = (id, Int64(k - 1))
    end
    id
end

"Copies the device bond state of block `blk` of env `id` into mat.general.intact."
function copy_intact!(mat, id::Int64, blk::Int64)
    ctx = context()
    intact = mat.general.intact
    rows = size(intact, 1)
    buf = zeros(UInt8, rows, size(intact, 2))
    check(ccall((:pd_block_family_export, LIB), Cint,
                (Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Ptr{UInt8}, Int64), ctx.h, id, blk, C_NULL, buf, rows), ctx)
    intact .= buf .!= 0
    nothing
end

"""
    force_density_T!(f, y, limits, mat, ::Type{Val{:hip}})

Counterpart of the CPU method (general_material.jl:296-369): out_of_loop!
passes, the bond loop with breaking, accumulated into f[:, limits[1]:limits[2]];
then the bond state is copied back into mat.general.intact.
"""
function PeriDyn.force_density_T!(f, y, limits, mat::PeriDyn.PeridynamicsMaterial, ::Type{Val{:hip}}; particles=nothing)
    ctx = context()
    if haskey(SEAM_BLOCK, mat)
        id, blk = SEAM_BLOCK[mat]
        yh = Array{Float64}(y)
        fh = Array{Float64}(f)
        GC.@preserve yh fh begin
            check(ccall((:pd_force_density, LIB), Cint,
                        (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Int64, Int64),
                        ctx.h, id, blk, yh, fh, limits[1], limits[2]), ctx)
        end
        f[:, limits[1]:limits[2]] .= fh[:, limits[1]:limits[2]]
    else
        id, blk = env_for(mat), Int64(0)
        yb = Array{Float64}(y[:, limits[1]:limits[2]])
        fb = zeros(Float64, size(yb))
        GC.@preserve yb fb begin
            check(ccall((:pd_force_density, LIB), Cint,
                        (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Int64, Int64),
                        ctx.h, id, 0, yb, fb, 1, size(yb, 2)), ctx)
        end
        f[:, limits[1]:limits[2]] .+= fb
    end
    copy_intact!(mat, id, blk)
    nothing
end

const SeamContact = Union{PeriDyn.NonLinearSpringContactModel12, PeriDyn.NonLinearSpringContactModel11}

seam_of(RM) = haskey(SEAM_CONTACT, RM) ? SEAM_CONTACT[RM] :
    error("PeriDynHIP: contact model $(RM.name) is not registered; call PeriDynHIP.register!(env) first")

"""
    short_range_contact!(y, f, type, bid, vol, den, RM)

For DEVICE[] == :hip: the contact accelerations of RM (RM12 contacts.jl:189-247,
RM11 :357-380) added into f, term by term in the single-threaded reference's
order (pd_contact_force), from the lists of the last update_contact_neighs!.
Other devices take the reference's own method.
"""
function PeriDyn.short_range_contact!(y, f, type, bid, vol, den, RM::SeamContact)
    if PeriDyn.DEVICE[] != :hip
        return invoke(PeriDyn.short_range_contact!,
                      Tuple{Any,Any,Any,Any,Any,Any,supertype(typeof(RM))}, y, f, type, bid, vol, den, RM)
    end
    ctx = context()
    id, k = seam_of(RM)
    yh = Array{Float64}(y)
    fh = Array{Float64}(f)
    GC.@preserve yh fh check(ccall((:pd_contact_force, LIB), Cint,
                                   (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Ptr{Float64}),
                                   ctx.h, id, k, yh, fh), ctx)
    f .= fh
    nothing
end

"""
    update_contact_neighs!(y, type, RM; max_part=nothing)

For DEVICE[] == :hip: RM's contact lists rebuilt on the device from y
(pd_contact_update) and copied into RM.neighs (pd_contact_neighs), as the
reference's method leaves them (contacts.jl:434-476 / 539-591).
"""
function PeriDyn.update_contact_neighs!(y, type, RM::SeamContact; max_part=nothing)
    if PeriDyn.DEVICE[] != :hip
        return invoke(PeriDyn.update_contact_neighs!, Tuple{Any,Any,supertype(typeof(RM))}, y, type, RM;
                      max_part=max_part)
    end
    ctx = context()
    id, k = seam_of(RM)
    yh = Array{Float64}(y)
    GC.@preserve yh check(ccall((:pd_contact_update, LIB), Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}),
                                ctx.h, id, k, yh), ctx)
    M = size(RM.neighs, 1)
    nb = zeros(Int64, M, size(RM.neighs, 2))
    check(ccall((:pd_contact_neighs, LIB), Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Int64),
                ctx.h, id, k, nb, M), ctx)
    RM.neighs .= nb
    nothing
end

"""
    upload_env(env; precision=PD_PREC_FP64) -> Int64

Uploads a whole Env once (environment.jl:166-235): its material blocks, its
short-range contacts (ContactModel12 pairs and ContactModel11 self contacts) and
its FixBC / ToFroBC / MoveBC conditions.  Returns the device env id; the state
then stays resident until `release(id)`.
"""
function upload_env(env; precision=PD_PREC_FP64)
    ctx = context()
    # the device counts steps from 0 at upload; a ToFroBC direction depends on
    # the absolute step (ToFroBC.jl:107-111)
    if env.time_step != 0 && any(bc -> !(bc isa PeriDyn.FixBC) && isfinite(bc.freq), env.boundary_conditions)
        error("upload_env: env.time_step = $(env.time_step) with a reversing ToFroBC; upload before stepping")
    end
    id = Ref{Int64}(0)
    check(ccall((:pd_env_create, LIB), Cint, (Ptr{Cvoid}, Float64, Int32, Ref{Int64}),
                ctx.h, env.dt, precision, id), ctx)
    keep = Any[]
    for mat in env.material_blocks
        gen, spc = mat.general, mat.specific
        nt = length(mat.type)
        arrs = (Array{Float64}(gen.x), Array{Float64}(gen.y), Array{Float64}(gen.velocity),
                Array{Float64}(gen.volume), Array{Int64}(gen.type), Array{Float64}(spc.density),
                Array{Float64}(spc.critical_stretch),
                hasproperty(spc, :bond_stiffness) ? Array{Float64}(spc.bond_stiffness) : zeros(nt, nt),
                hasproperty(spc, :shear_modulus) ? Array{Float64}(spc.bulk_modulus) : zeros(nt, nt),
                hasproperty(spc, :shear_modulus) ? Array{Float64}(spc.shear_modulus) : zeros(nt, nt),
                hasproperty(spc, :ψ) ? Array{Float64}(spc.ψ) : zeros(nt, nt))
        push!(keep, arrs)
        x, y, v, vol, typ, dens, s0, c, K, G, psi = arrs
        d = BlockDesc(kind_of(mat), Int32(nt), size(x, 2), first(mat.type),
                      pointer(x), pointer(y), pointer(v), pointer(vol), pointer(typ),
                      gen.horizon, gen.particle_size, family_rows(gen),
                      pointer(dens), pointer(s0), pointer(c), pointer(K), pointer(G), pointer(psi),
                      Int32(all(gen.deformed)), Int32(0))
        GC.@preserve arrs check(ccall((:pd_env_add_block, LIB), Cint,
                                      (Ptr{Cvoid}, Int64, Ref{BlockDesc}, Ptr{Int64}), ctx.h, id[], d, C_NULL), ctx)
    end
    for RM in env.short_range_contact
        # ContactModel11 (contacts.jl:105-119): one block, no second type list
        selfc = RM isa PeriDyn.ContactModel11
        t1 = selfc ? collect(Int64, RM.type) : collect(Int64, RM.pair[1])
        t2 = selfc ? Int64[] : collect(Int64, RM.pair[2])
        push!(keep, (t1, t2))
        e = Float64(RM.exponent)        # stored as Float64 (nonlinearspring.jl:16,33)
        cd = ContactDesc(pointer(t1), length(t1), selfc ? Ptr{Int64}(C_NULL) : pointer(t2), length(t2), e,
                         Int32(isinteger(e) && abs(e) <= 4096), Int32(0), RM.stifness, RM.equi_dist,
                         RM.distance, RM.max_neighs)
        GC.@preserve t1 t2 check(ccall((:pd_env_add_contact, LIB), Cint,
                                       (Ptr{Cvoid}, Int64, Ref{ContactDesc}), ctx.h, id[], cd), ctx)
    end
    for bc in env.boundary_conditions
        mask = UInt8.(Array(bc.bool))
        push!(keep, mask)
        # onlyatstart and dims of the BC constructors (FixBC.jl:46, ToFroBC.jl:51); the
        # library refuses a dims subset, which the reference's update cannot take either
        dims = Array(bc.dims)
        opts = Int32((bc.onlyatstart ? PD_BC_ONLYATSTART : 0) |
                     ((Int(dims[1] != 0) | (Int(dims[2] != 0) << 1) | (Int(dims[3] != 0) << 2)) << 8))
        if bc isa PeriDyn.FixBC
            bd = BcDesc(PD_BC_FIX, opts, pointer(mask), (0.0, 0.0, 0.0), Inf, 0)
        else
            r = Array(bc.rate)
            bd = BcDesc(PD_BC_TOFRO, opts, pointer(mask), (r[1], r[2], r[3]), Float64(bc.freq), bc.applyafter)
        end
        GC.@preserve mask check(ccall((:pd_env_add_bc, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{BcDesc}),
                                      ctx.h, id[], bd), ctx)
    end
    check(ccall((:pd_env_commit, LIB), Cint, (Ptr{Cvoid}, Int64), ctx.h, id[]), ctx)
    id[]
end

"Copies the resident y, v, f of device env `id` into env; env.time_step = t0 + device steps."
function download!(env, id::Int64, t0::Integer)
    ctx = context()
    y = similar(env.y, Float64); v = similar(env.v, Float64); f = similar(env.f, Float64)
    check(ccall((:pd_env_download, LIB), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ctx.h, id, y, v, f), ctx)
    env.y .= y; env.v .= v; env.f .= f
    ts = Ref{Int64}(0)
    check(ccall((:pd_env_time_step, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{Int64}), ctx.h, id, ts), ctx)
    env.time_step = t0 + ts[]
    env
end

release(id::Int64) = check(ccall((:pd_env_destroy, LIB), Cint, (Ptr{Cvoid}, Int64), context().h, id), context())

"""
    run_resident!(env, N; neigh_update_freq=1, precision=PD_PREC_FP64, average_prop_freq=0)

Uploads the Env, runs N velocity_verlet_step!s on the device with run!'s
contact-update cadence (solvers.jl:97-187), copies y, v, f and time_step back.
With average_prop_freq > 0 it also returns the printdata rows that run! logs
(time, momentum, force sum, mean damage; environment.jl:308-320) after step 1
and every average_prop_freq-th step, as a (rows, 8) matrix.
"""
function run_resident!(env, N::Integer; neigh_update_freq::Integer=1, precision=PD_PREC_FP64,
                       average_prop_freq::Integer=0)
    ctx = context()
    t0 = env.time_step
    id = upload_env(env; precision=precision)
    ids = [id]
    rows = nothing
    if average_prop_freq > 0
        nrow = ccall((:pd_log_rows, LIB), Int64, (Int64, Int64), N, average_prop_freq)
        log = zeros(Float64, 8, nrow)
        check(ccall((:pd_run_logged, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}, Int64, Int64, Int64, Int64, Ptr{Float64}),
                    ctx.h, ids, 1, N, neigh_update_freq, average_prop_freq, log), ctx)
        rows = permutedims(log)
        rows[:, 1] .+= t0 * env.dt
    else
        check(ccall((:pd_run, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}, Int64, Int64, Int64),
                    ctx.h, ids, 1, N, neigh_update_freq), ctx)
    end
    download!(env, id, t0)
    release(id)
    average_prop_freq > 0 ? (env, rows) : env
end

"""
    minimize_resident!(env, solver::PeriDyn.QSDrag; precision=PD_PREC_FP64)

QSDrag minimize! (minimize.jl:105-231) on the device: BCs once, then drag
iterations of update_acc! until max|f| < f_tol and max|dy| < x_tol or max_iter.
Returns (env, iterations).
"""
function minimize_resident!(env, solver; precision=PD_PREC_FP64)
    ctx = context()
    t0 = env.time_step
    id = upload_env(env; precision=precision)
    it = Ref{Int64}(0); fm = Ref{Float64}(0); xm = Ref{Float64}(0)
    check(ccall((:pd_env_minimize, LIB), Cint,
                (Ptr{Cvoid}, Int64, Float64, Float64, Int64, Float64, Float64, Ref{Int64}, Ref{Float64}, Ref{Float64}),
                ctx.h, id, solver.step_size, solver.drag, solver.max_iter, solver.x_tol, solver.f_tol, it, fm, xm), ctx)
    download!(env, id, t0)
    release(id)
    env, it[]
end

"Resident state of device env `id` as bytes (pd_env_state_save); `restore!` loads it back."
function checkpoint(id::Int64)
    ctx = context()
    n = Ref{Int64}(0)
    check(ccall((:pd_env_state_size, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{Int64}), ctx.h, id, n), ctx)
    buf = Vector{UInt8}(undef, n[])
    check(ccall((:pd_env_state_save, LIB), Cint, (Ptr{Cvoid}, Int64, Ptr{UInt8}, Int64), ctx.h, id, buf, n[]), ctx)
    buf
end

restore!(id::Int64, buf::Vector{UInt8}) =
    check(ccall((:pd_env_state_load, LIB), Cint, (Ptr{Cvoid}, Int64, Ptr{UInt8}, Int64),
                context().h, id, buf, length(buf)), context())

end # module
