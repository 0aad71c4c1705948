"""
    PeriDynHIP

Binds PeriDyn's explicit-dynamics hot path to libperidyn_hip.so (MI355X, gfx950)
through `ccall`.  It adds `Val{:hip}` methods at the reference's own dispatch
seams:

  * force_density_T!(f, y, limits, mat, ::Type{Val{:hip}})   (general_material.jl:274-296)
  * short_range_contact!(y, f, type, bid, vol, den, RM, ::Type{Val{:hip}}) (contacts.jl:210)
  * update_contact_neighs!(y, type, RM::ContactModel12; device=:hip)       (contacts.jl:434)
  * run_resident!(env, N; neigh_update_freq)  — the resident, fused run loop
    (run!'s physics cadence, solvers.jl:97-187, without file output)

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
    kind::Int32; reserved::Int32; mask::Ptr{UInt8}; rate::NTuple{3,Float64}
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
                          gen.horizon, gen.particle_size, size(gen.family.x, 1) * 2 + 1,
                          pointer(dens), pointer(s0), pointer(c), pointer(K), pointer(G), pointer(psi),
                          Int32(all(gen.deformed)), Int32(0))
            check(ccall((:pd_env_add_block, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{BlockDesc}, Ptr{Int64}),
                        ctx.h, id[], d, C_NULL), ctx)
            check(ccall((:pd_env_commit, LIB), Cint, (Ptr{Cvoid}, Int64), ctx.h, id[]), ctx)
        end
        id[]
    end
end

"""
    force_density_T!(f, y, limits, mat, ::Type{Val{:hip}})

Counterpart of the CPU method (general_material.jl:296-369): out_of_loop!
passes, the bond loop with breaking, accumulated into f[:, limits[1]:limits[2]].
Bond state, theta and edp stay resident on the device.
"""
function PeriDyn.force_density_T!(f, y, limits, mat::PeriDyn.PeridynamicsMaterial, ::Type{Val{:hip}}; particles=nothing)
    ctx = context()
    id = env_for(mat)
    yb = Array{Float64}(y[:, limits[1]:limits[2]])
    fb = zeros(Float64, size(yb))
    GC.@preserve yb fb begin
        check(ccall((:pd_force_density, LIB), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Int64, Int64),
                    ctx.h, id, 0, yb, fb, 1, size(yb, 2)), ctx)
    end
    f[:, limits[1]:limits[2]] .+= fb
    nothing
end

"""
    run_resident!(env, N; neigh_update_freq=1, precision=PD_PREC_FP64)

Uploads the whole Env once (blocks, contacts, FixBC/ToFroBC/MoveBC) and runs N
velocity_verlet_step!s on the device with run!'s contact-update cadence; y, v
and f are copied back into env at the end.
"""
function run_resident!(env, N::Integer; neigh_update_freq::Integer=1, precision=PD_PREC_FP64)
    ctx = context()
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
                      gen.horizon, gen.particle_size, 1 << 20,
                      pointer(dens), pointer(s0), pointer(c), pointer(K), pointer(G), pointer(psi),
                      Int32(all(gen.deformed)), Int32(0))
        GC.@preserve arrs check(ccall((:pd_env_add_block, LIB), Cint,
                                      (Ptr{Cvoid}, Int64, Ref{BlockDesc}, Ptr{Int64}), ctx.h, id[], d, C_NULL), ctx)
    end
    for RM in env.short_range_contact
        t1 = collect(Int64, RM.pair[1]); t2 = collect(Int64, RM.pair[2])
        push!(keep, (t1, t2))
        cd = ContactDesc(pointer(t1), length(t1), pointer(t2), length(t2), Float64(RM.exponent),
                         Int32(RM.exponent isa Integer), Int32(0), RM.stifness, RM.equi_dist,
                         RM.distance, RM.max_neighs)
        GC.@preserve t1 t2 check(ccall((:pd_env_add_contact, LIB), Cint,
                                       (Ptr{Cvoid}, Int64, Ref{ContactDesc}), ctx.h, id[], cd), ctx)
    end
    for bc in env.boundary_conditions
        mask = UInt8.(Array(bc.bool))
        push!(keep, mask)
        if bc isa PeriDyn.FixBC
            bd = BcDesc(PD_BC_FIX, 0, pointer(mask), (0.0, 0.0, 0.0), Inf, 0)
        else
            r = Array(bc.rate)
            bd = BcDesc(PD_BC_TOFRO, 0, pointer(mask), (r[1], r[2], r[3]), Float64(bc.freq), bc.applyafter)
        end
        GC.@preserve mask check(ccall((:pd_env_add_bc, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{BcDesc}),
                                      ctx.h, id[], bd), ctx)
    end
    check(ccall((:pd_env_commit, LIB), Cint, (Ptr{Cvoid}, Int64), ctx.h, id[]), ctx)
    ids = [id[]]
    check(ccall((:pd_run, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}, Int64, Int64, Int64),
                ctx.h, ids, 1, N, neigh_update_freq), ctx)
    y = similar(env.y, Float64); v = similar(env.v, Float64); f = similar(env.f, Float64)
    check(ccall((:pd_env_download, LIB), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ctx.h, id[], y, v, f), ctx)
    env.y .= y; env.v .= v; env.f .= f
    env.time_step += N
    check(ccall((:pd_env_destroy, LIB), Cint, (Ptr{Cvoid}, Int64), ctx.h, id[]), ctx)
    env
end

end # module
