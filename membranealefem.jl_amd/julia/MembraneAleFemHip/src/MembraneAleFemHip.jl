# MembraneAleFemHip: runs MembraneAleFem.jl's per-time-step hot path
# (calc_r_K + K \ r inside time_step!) on an MI355X through libmafem's C ABI
# (include/mafem.h).  Loading this package after MembraneAleFem replaces
# `time_step!`, `calc_r_K` and `calc_pull_force`; `solve`, `restart`,
# `prepare_input` and `run_analysis!` stay the reference's own.
#
#   using MembraneAleFem, MembraneAleFemHip
#   MembraneAleFemHip.enable!()            # route time_step!/calc_r_K to the GPU
#   solve(p; pull_speed=0.5, Δts=..., t0=0, t0_id=0, out_path=".", out_file="out.txt")
module MembraneAleFemHip

using SparseArrays
import MembraneAleFem
const MAF = MembraneAleFem

const libmafem = get(ENV, "MAFEM_LIB",
                     joinpath(@__DIR__, "..", "..", "..", "libmafem.so"))

const MAFEM_OK = Cint(0)
const MAFEM_NOT_CONVERGED = Cint(1)

# include/mafem.h: mafem_mesh_desc
struct MeshDesc
  num1el::Int64; num2el::Int64; numel::Int64
  num1np::Int64; num2np::Int64; numnp::Int64
  ndf::Int64; nmdf::Int64
  IX::Ptr{Int64}; ID::Ptr{Int64}
  v_order::NTuple{3,Int32}; m_order::NTuple{3,Int32}
  lam_order::Int32; p_order::Int32
  area_n_uel::Int64; area_uel_ids::Ptr{Int64}; area_ufns::Ptr{Float64}
  bdry_n_el::NTuple{4,Int64}; bdry_elems::NTuple{4,Ptr{Int64}}
  bdry_n_uel::NTuple{4,Int64}; bdry_uel_ids::NTuple{4,Ptr{Int64}}
  bdry_ufns::NTuple{4,Ptr{Float64}}
  n_neu::Int64; neu_bdry::Ptr{Int32}; neu_type::Ptr{Int32}; neu_val::Ptr{Float64}
  crnr_ufns::Ptr{Float64}
end

# include/mafem.h: mafem_params
struct ParamsDesc
  motion::Int32; scenario::Int32
  kb::Float64; kg::Float64; zv::Float64; pn::Float64; adb::Float64; am::Float64
  ek::Float64; enr::Float64; bend_tm::Float64
end

mutable struct Context
  h::Ptr{Cvoid}
  mesh::MAF.Mesh
  keep::Vector{Any}          # arrays the C side saw at create time
  nnz::Int64
  function Context(mesh::MAF.Mesh, p::MAF.Params; bend_tm=1.0)
    bdrys = (MAF.BOTTOM, MAF.RIGHT, MAF.TOP, MAF.LEFT)
    dofs = mesh.dofs
    ord(u) = Int32(get(dofs, u, 0))
    neu_b = Int32[Int32(b) for (b, _, _) in mesh.inh_neu_bcs]
    neu_t = Int32[Int32(t) for (_, t, _) in mesh.inh_neu_bcs]
    neu_v = Float64[v for (_, _, v) in mesh.inh_neu_bcs]
    belems = [mesh.bdry_elems[b] for b in bdrys]
    bfns = [mesh.bdry_gp_fns[b] for b in bdrys]
    crnr = [mesh.crnr_gp_fns[c]
            for c in (MAF.BOTTOM_LEFT, MAF.BOTTOM_RIGHT, MAF.TOP_LEFT, MAF.TOP_RIGHT)]
    keep = Any[mesh.IX, mesh.ID, mesh.area_gp_fns.uel_ids, mesh.area_gp_fns.ufns,
               belems..., [f.uel_ids for f in bfns]..., [f.ufns for f in bfns]...,
               neu_b, neu_t, neu_v, crnr]
    GC.@preserve keep begin
      d = MeshDesc(mesh.num1el, mesh.num2el, mesh.numel, mesh.num1np, mesh.num2np, mesh.numnp,
                   mesh.ndf, mesh.nmdf, pointer(mesh.IX), pointer(mesh.ID),
                   (ord(MAF.Dof.vx), ord(MAF.Dof.vy), ord(MAF.Dof.vz)),
                   (ord(MAF.Dof.vmx), ord(MAF.Dof.vmy), ord(MAF.Dof.vmz)),
                   ord(MAF.Dof.λ), ord(MAF.Dof.pm),
                   size(mesh.area_gp_fns.ufns, 1), pointer(mesh.area_gp_fns.uel_ids),
                   Ptr{Float64}(pointer(mesh.area_gp_fns.ufns)),
                   Tuple(Int64(length(e)) for e in belems), Tuple(pointer(e) for e in belems),
                   Tuple(Int64(size(f.ufns, 1)) for f in bfns),
                   Tuple(pointer(f.uel_ids) for f in bfns),
                   Tuple(Ptr{Float64}(pointer(f.ufns)) for f in bfns),
                   length(neu_b), pointer(neu_b), pointer(neu_t), pointer(neu_v),
                   Ptr{Float64}(pointer(crnr)))
      q = ParamsDesc(Int32(p.motion), Int32(p.scenario), p.kb, p.kg, p.ζv, p.pn, p.αdb, p.αm,
                     p.εk, p.εnr, bend_tm)
      hr = Ref{Ptr{Cvoid}}(C_NULL)
      rc = ccall((:mafem_create, libmafem), Cint,
                 (Ref{Ptr{Cvoid}}, Ref{MeshDesc}, Ref{ParamsDesc}), hr, d, q)
      rc == MAFEM_OK || error("mafem_create: " *
                              unsafe_string(ccall((:mafem_last_error, libmafem), Cstring,
                                                  (Ptr{Cvoid},), C_NULL)))
    end
    nnz = Ref{Int64}(0)
    ccall((:mafem_nnz, libmafem), Cint, (Ptr{Cvoid}, Ref{Int64}), hr[], nnz)
    ctx = new(hr[], mesh, keep, nnz[])
    finalizer(c -> ccall((:mafem_destroy, libmafem), Cvoid, (Ptr{Cvoid},), c.h), ctx)
    return ctx
  end
end

check(ctx, rc) = rc >= 0 ? rc :
  error("libmafem: " * unsafe_string(ccall((:mafem_last_error, libmafem), Cstring,
                                           (Ptr{Cvoid},), ctx.h)))

const CONTEXTS = IdDict{MAF.Mesh,Context}()
context(mesh, p; args...) =
  get!(() -> Context(mesh, p; bend_tm=get(args, :bend_tm, 1.0)), CONTEXTS, mesh)

set_state!(ctx, xms, cps) = check(ctx, ccall((:mafem_set_state, libmafem), Cint,
                                             (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                                             ctx.h, xms, cps))
get_state!(ctx, xms, cps) = check(ctx, ccall((:mafem_get_state, libmafem), Cint,
                                             (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                                             ctx.h, xms, cps))

"calc_r_K on the GPU: (r_gl, K_gl::SparseMatrixCSC) as FiniteElement.jl:75-200 returns."
function calc_r_K(mesh, xms, cps, time, Δt, p; args...)
  ctx = context(mesh, p; args...)
  set_state!(ctx, xms, cps)
  n = mesh.nmdf
  r = zeros(n); colptr = zeros(Int64, n + 1); rowval = zeros(Int64, ctx.nnz)
  nzval = zeros(ctx.nnz)
  check(ctx, ccall((:mafem_assemble, libmafem), Cint,
                   (Ptr{Cvoid}, Float64, Float64, Ptr{Float64}, Ptr{Int64}, Ptr{Int64},
                    Ptr{Float64}), ctx.h, time, Δt, r, colptr, rowval, nzval))
  return r, SparseMatrixCSC(n, n, colptr, rowval, nzval)
end

"time_step! on the GPU, with the reference's output lines and AssertionError."
function time_step!(mesh, xms, cps, time, Δt, p; args...)
  ctx = context(mesh, p; args...)
  set_state!(ctx, xms, cps)
  εs = zeros(14); iters = Ref{Int32}(0)
  rc = check(ctx, ccall((:mafem_time_step, libmafem), Cint,
                        (Ptr{Cvoid}, Float64, Float64, Int32, Ref{Int32}, Ptr{Float64}),
                        ctx.h, time, Δt, Int32(14), iters, εs))
  get_state!(ctx, xms, cps)
  if p.output
    open(args[:out_path] * "/" * args[:out_file], "a") do io
      for k = 1:iters[]
        println(io, "--> iteration ", k, ": ε = ", εs[k])
      end
    end
  end
  @assert rc != MAFEM_NOT_CONVERGED "did not reach Newton--Raphson tolerance"
  return Nothing
end

"calc_pull_force on the GPU (PullForce.jl:61-80); the device builds the same
adjacent elements and node maps as get_adj_maps with poly = 2."
function calc_pull_force(mesh, xms, cps, adj_el_ids, adj_node_map, p)
  ctx = context(mesh, p)
  set_state!(ctx, xms, cps)
  x = zeros(3); f = zeros(3)
  check(ctx, ccall((:mafem_pull_force, libmafem), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), ctx.h, x, f))
  return f
end

"generate_output (Output.jl:30-118) on the GPU, for a mesh that has a context."
function generate_output(mesh, xms, cps)
  ctx = CONTEXTS[mesh]
  set_state!(ctx, xms, cps)
  n1 = mesh.num1el * 3 + 2; n2 = mesh.num2el * 3 + 2
  xout = zeros(n1, n2, 3); uout = zeros(n1, n2, mesh.ndf)
  check(ctx, ccall((:mafem_generate_output, libmafem), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), ctx.h, xout, uout))
  return xout, uout
end

"Route MembraneAleFem's time_step!/calc_r_K/calc_pull_force to the GPU (call after loading)."
function enable!()
  @eval MAF function time_step!(mesh::Mesh, xms::Matrix{Float64}, cps::Matrix{Float64},
                                time::Float64, Δt::Float64, p::Params; args...)
    return $(time_step!)(mesh, xms, cps, time, Δt, p; args...)
  end
  @eval MAF function calc_r_K(mesh::Mesh, xms::Matrix{Float64}, cps::Matrix{Float64},
                              time::Float64, Δt::Float64, p::Params; args...)
    return $(calc_r_K)(mesh, xms, cps, time, Δt, p; args...)
  end
  @eval MAF function calc_pull_force(mesh::Mesh, xms::Matrix{Float64}, cps::Matrix{Float64},
                                     adj_el_ids::Vector{Int64},
                                     adj_node_map::Vector{Tuple{Vector{Int64},Vector{Int64}}},
                                     p::Params)::Vector{Float64}
    return $(calc_pull_force)(mesh, xms, cps, adj_el_ids, adj_node_map, p)
  end
  return nothing
end

end # module
