# SplatHIP.jl -- the binding a SPLAT maintainer adds to call the MI355X light-map path.
#
# New hitter/tracer types plug into the reference's dispatch (src/alg_types.jl) so that
# run_evolution!/trace! pick the gfx950 kernels, through `ccall` into libsplat_hip.so
# (include/splat_hip.h). Device arrays are AMDGPU.jl ROCArrays holding the reference's own isbits
# types (ADRay, FTri, Tuple{Int32,Tri}), passed by pointer with zero copies; the layouts match
# byte for byte (136/120/52 bytes). Not run in this repository: there is no Julia toolchain here.
module SplatHIP

using AMDGPU

const LIB = joinpath(@__DIR__, "..", "splat_amd", "libsplat_hip.so")

struct SplatError <: Exception
    code::Int32
    where::String
end
Base.showerror(io::IO, e::SplatError) =
    print(io, e.where, ": ", unsafe_string(ccall((:splat_strerror, LIB), Cstring, (Int32,), e.code)))
check(code, where) = code == 0 || throw(SplatError(code, where))

stream_ptr() = AMDGPU.stream().stream  # the task-local HIP stream handle

# ---- C structs (include/splat_hip.h) -----------------------------------------------------------
struct Optics
    ior_model::Int32
    cauchy_a::Float32
    cauchy_b::Float32
    compat::Int32
end
Optics(; model = 0, a = 1.5f0, b = 0.004f0, compat = 3) = Optics(model, a, b, compat)  # 0 = the reference's glass()

mutable struct HitSet
    kind::Int32; n_tris::Int32; n_bv::Int32; max_members::Int32
    total_members::Int64
    d_recs::Ptr{Cvoid}; d_boxes::Ptr{Cvoid}; d_ranges::Ptr{Cvoid}; d_groups::Ptr{Cvoid}
    group_size::Int32; n_groups::Int32; tune_waves::Int32; tree_levels::Int32
    top_bvs::Int32; top_recs::Int32
    d_stats::Ptr{UInt64}
    HitSet() = new(0, 0, 0, 0, 0, C_NULL, C_NULL, C_NULL, C_NULL, 0, 0, 0, 0, 0, 0, C_NULL)
end

struct RunParams
    depth::Int32; first_diffuse::Int32; force_rand::Float32; _pad::Int32
    seed::UInt64
    optics::Optics
    hit_events::Ptr{Ptr{Cvoid}}; n_hit_events::Int32; _pad2::Int32
end

# ---- hitters (subtypes of the reference's AbstractHitter) -------------------------------------
struct HIPBruteForceHitter <: AbstractHitter
    hs::HitSet
    ws::ROCArray{UInt8}
end
function HIPBruteForceHitter(n_tris::ROCArray{Tuple{Int32,Tri}})
    N = length(n_tris)
    ws = ROCArray{UInt8}(undef, ccall((:splat_hitset_bf_bytes, LIB), Csize_t, (Int32,), N))
    hs = HitSet()
    check(ccall((:splat_hitset_prepare_bf, LIB), Int32,
                (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Csize_t, Ref{HitSet}, Ptr{Cvoid}),
                pointer(n_tris), N, pointer(ws), length(ws), hs, stream_ptr()), "prepare_bf")
    HIPBruteForceHitter(hs, ws)
end

struct HIPBVHitter <: AbstractHitter
    hs::HitSet
    ws::ROCArray{UInt8}
end
# bvs/memberships as returned by bv_partition (src/bv.jl:92-121); packed like pack_bv_tris
function HIPBVHitter(tris::ROCArray{FTri}, bvs, memberships; group_size = 0, n_top = 0)
    counts, bv_tris = pack_bv_tris(Array, Array(tris), bvs, memberships)  # host copies
    n_bv = length(bvs)
    total = sum(counts)
    ws = ROCArray{UInt8}(undef, ccall((:splat_hitset_bv_bytes, LIB), Csize_t, (Int32, Int64), n_bv, total))
    hs = HitSet()
    check(ccall((:splat_hitset_prepare_bv, LIB), Int32,
                (Ptr{Cvoid}, Int32, Int32, Ptr{Int64}, Ptr{Int64}, Int64, Int32, Int32, Ptr{Cvoid}, Csize_t, Ref{HitSet},
                 Ptr{Cvoid}),
                pointer(tris), length(tris), n_bv, counts, bv_tris, size(bv_tris, 2), group_size, n_top,
                pointer(ws), length(ws), hs, stream_ptr()), "prepare_bv")
    HIPBVHitter(hs, ws)
end

# The tree hitter's BVs built on the device from the triangles alone (splat_hitset_build_bv) in
# place of bv_partition + pack_bv_tris on the host: ~0.4 ms for 10k triangles, so an animated
# scene (src/scenes/splash_scene.jl:41) can rebuild every frame with rebuild!.
function HIPBVHitter(tris::ROCArray{FTri}; depth = max(1, round(Int, log2(max(length(tris) - 1, 2) / 8))),
                     big_factor = 16f0, max_big = 32)
    nbytes = ccall((:splat_hitset_build_bv_bytes, LIB), Csize_t, (Int32, Int32), length(tris), depth)
    nbytes == 0 && throw(SplatError(Int32(-4), "build_bv"))
    h = HIPBVHitter(HitSet(), ROCArray{UInt8}(undef, nbytes))
    rebuild!(h, tris; depth, big_factor, max_big)
end
function rebuild!(h::HIPBVHitter, tris::ROCArray{FTri}; depth, big_factor = 16f0, max_big = 32)
    check(ccall((:splat_hitset_build_bv, LIB), Int32,
                (Ptr{Cvoid}, Int32, Int32, Float32, Int32, Ptr{Cvoid}, Csize_t, Ref{HitSet}, Ptr{Cvoid}),
                pointer(tris), length(tris), depth, big_factor, max_big, pointer(h.ws), length(h.ws), h.hs,
                stream_ptr()), "build_bv")
    h
end

# device workspace of splat_next_hit (the packed (t, index) key per ray), one per tracer
const HIT_WS = IdDict{Any,ROCArray{UInt8}}()
function hit_workspace(tracer, R)
    need = ccall((:splat_next_hit_ws_bytes, LIB), Csize_t, (Int64,), R)
    ws = get(HIT_WS, tracer, nothing)
    if ws === nothing || length(ws) < need
        ws = HIT_WS[tracer] = ROCArray{UInt8}(undef, need)
    end
    return ws
end

# next_hit!(tracer, hitter, rays, n_tris) -- src/hitters.jl:24-30 semantics
function next_hit!(tracer, hitter::Union{HIPBruteForceHitter,HIPBVHitter}, rays::ROCArray{ADRay}, n_tris)
    ws = hit_workspace(tracer, length(rays))  # ROCArray{UInt8} of splat_next_hit_ws_bytes(R), kept by the tracer
    check(ccall((:splat_next_hit, LIB), Int32,
                (Ptr{Cvoid}, Int64, Ptr{Int64}, Ref{HitSet}, Ptr{Int32}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                pointer(rays), length(rays), C_NULL, hitter.hs, pointer(tracer.hit_idx), pointer(ws), length(ws),
                stream_ptr()), "next_hit")
end

# partition!(rays, swap; by = r.status != ACTIVE) -- src/partition.jl:122-132, stable
function partition!(vals::ROCArray{ADRay}, swap::ROCArray{ADRay}; kwargs...)
    R = length(vals)
    ws = ROCArray{UInt8}(undef, ccall((:splat_partition_ws_bytes, LIB), Csize_t, (Int64,), R))
    d = ROCArray{Int64}(undef, 1)
    h = Ref{Int64}(0)
    check(ccall((:splat_partition_active, LIB), Int32,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Int64}, Ref{Int64}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                pointer(vals), pointer(swap), R, pointer(d), h, pointer(ws), length(ws), stream_ptr()), "partition")
    return h[]
end

# run_evolution!(hitter, ::ExperimentalTracer; ...) -- src/tracer.jl:233-282, device-resident counts
function run_evolution!(hitter::Union{HIPBruteForceHitter,HIPBVHitter}, tracer::ExperimentalTracer;
                        depth, first_diffuse, tris, rays, force_rand = nothing, seed = 0, optics = Optics(),
                        reorder = false, kwargs...)
    R = length(rays)
    ws = ROCArray{UInt8}(undef, ccall((:splat_run_ws_bytes, LIB), Csize_t, (Int64,), R))
    live = ROCArray{Int64}(undef, 1)
    p = RunParams(depth, first_diffuse, isnothing(force_rand) ? NaN32 : Float32(force_rand), 0, seed, optics,
                  C_NULL, 0, 0)
    check(ccall((:splat_run_evolution, LIB), Int32,
                (Ptr{Cvoid}, Int64, Ref{HitSet}, Ptr{Cvoid}, Int32, Ptr{Int32}, Ref{RunParams}, Ptr{Cvoid}, Csize_t,
                 Ptr{Int64}, Ptr{Cvoid}),
                pointer(rays), R, hitter.hs, pointer(tris), length(tris), pointer(tracer.hit_idx), p,
                pointer(ws), length(ws), pointer(live), stream_ptr()), "run_evolution")
    reorder && reorder!(tracer, rays)  # the backward camera pass (src/tracer.jl:270-278)
end


# run_evolution!(...; reorder = true)'s reorder (src/tracer.jl:270-278)
function reorder!(tracer, rays::ROCArray{ADRay})
    R = length(rays)
    ws = ROCArray{UInt8}(undef, ccall((:splat_reorder_ws_bytes, LIB), Csize_t, (Int64,), R))
    check(ccall((:splat_reorder_by_dest, LIB), Int32,
                (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                pointer(rays), R, pointer(tracer.hit_idx), pointer(ws), length(ws), stream_ptr()), "reorder")
end

# continuum_shade!(::ExperimentalImager2; ...) (src/ray_imagers.jl:440-503): lm holds the spectrum,
# the sub-pixel offsets and (weights) the retina factors; RGB3 is (3, n_pixels) Float32
function continuum_shade!(tracer, rays::ROCArray{ADRay}, tris::ROCArray{FTri}, lm::Ref, tex::ROCArray{Float32},
                          skydome::Union{Nothing,ROCArray{Float32,3}}, grid_h, RGB3::ROCArray{Float32})
    sky = skydome === nothing ? C_NULL : pointer(skydome)
    sw = skydome === nothing ? 0 : size(skydome, 1)
    sh = skydome === nothing ? 0 : size(skydome, 2)
    check(ccall((:splat_shade, LIB), Int32,
                (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Cvoid}, Int32, Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Int32, Int32,
                 Int64, Ptr{Float32}, Ptr{Cvoid}),
                pointer(rays), length(rays), pointer(tracer.hit_idx), pointer(tris), length(tris), lm, pointer(tex),
                sky, sw, sh, grid_h, pointer(RGB3), stream_ptr()), "shade")
end

# rays_from_lights(lights, upscale) (src/ray_generators.jl:64-113) on the device. The reference's
# Vector{RectLight} is passed as is (64-byte isbits records); rand(Float32)'s cell offsets come
# from `seed` (lambda_mode 0 = every ray at lambda_min, 1 = uniform in [lambda_min, lambda_max)).
struct LightRays
    upscale::Int32; lambda_mode::Int32; lambda_min::Float32; lambda_max::Float32
    jitter::Float32; _pad::Int32
    seed::UInt64
end
LightRays(upscale; lambda_mode = 0, lambda_min = 550f0, lambda_max = 550f0, jitter = 1f0, seed = 0) =
    LightRays(upscale, lambda_mode, lambda_min, lambda_max, jitter, 0, seed)

function rays_from_lights(lights::Vector{RectLight}, upscale; kw...)
    p = LightRays(upscale; kw...)
    R = ccall((:splat_lights_ray_count, LIB), Int64, (Ptr{RectLight}, Int32, Int32), lights, length(lights), upscale)
    R < 0 && throw(SplatError(Int32(R), "rays_from_lights"))
    rays = ROCArray{ADRay}(undef, R)
    check(ccall((:splat_raygen_lights, LIB), Int32,
                (Ptr{RectLight}, Int32, Ref{LightRays}, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                lights, length(lights), p, pointer(rays), R, stream_ptr()), "rays_from_lights")
    return rays
end

end # module
