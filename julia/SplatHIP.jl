# SplatHIP.jl -- the binding a SPLAT maintainer adds to call the MI355X light-map path.
#
# Load it after the reference (`include("tracer.jl"); include("SplatHIP.jl")`): it defines new
# hitter and tracer types under the reference's abstract types (src/alg_types.jl) and new methods
# of its generic functions (next_hit!, partition!, run_evolution!, continuum_light_map2!,
# continuum_shade!, trace!) that `ccall` libsplat_hip.so (include/splat_hip.h). Device arrays are
# AMDGPU.jl ROCArrays holding the reference's own isbits types (ADRay 136 B, FTri 120 B,
# Tuple{Int32,Tri} 52 B), passed by pointer with zero copies. Every workspace lives in the tracer
# or hitter and is reused: after the first frame of a given size no call allocates device memory.
# Not run in this repository: there is no Julia toolchain here.
module SplatHIP

using AMDGPU
import Main: next_hit!, partition!, run_evolution!, continuum_light_map2!, continuum_shade!, trace!
using Main: AbstractHitter, AbstractTracer, ADRay, FTri, Tri, RectLight, Cam, ExperimentalImager2, upres

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
RunParams(depth, first_diffuse, force_rand, seed, optics) =
    RunParams(depth, first_diffuse, isnothing(force_rand) ? NaN32 : Float32(force_rand), 0, seed, optics, C_NULL, 0, 0)

const MAX_LAMBDA, MAX_DELTA, MAX_CURVE = 64, 8, 24
const WEIGHTS_IDENTITY, WEIGHTS_TABLE = Int32(0), Int32(1)

# splat_lightmap_params_t: the splat's wavelength grid, sub-pixel offsets and weighting
struct LightMapParams
    n_lambda::Int32; n_delta::Int32; channels::Int32; weight_mode::Int32
    width::Int32; height::Int32
    spectrum::NTuple{MAX_LAMBDA,Float32}
    delta::NTuple{MAX_DELTA,Float32}
    weights::NTuple{3 * MAX_LAMBDA,Float32}
    curve_n::NTuple{3,Int32}
    curve_x::NTuple{3 * MAX_CURVE,Float32}
    curve_y::NTuple{3 * MAX_CURVE,Float32}
end
padded(v, n) = ntuple(i -> i <= length(v) ? Float32(v[i]) : 0f0, n)
# continuum_light_map2!'s tex[i, j, k] += cos for lambda = spectrum[k] (one plane per wavelength);
# retina_factor (3 x n_lambda, src/rgb_spectrum.jl:135-148) is kept in `weights` for the shading
function LightMapParams(width, height, spectrum, δ; retina_factor = nothing)
    n = length(spectrum)
    (1 <= n <= MAX_LAMBDA && 1 <= length(δ) <= MAX_DELTA) || throw(SplatError(Int32(-4), "light-map parameters"))
    w = zeros(Float32, 3 * MAX_LAMBDA)
    if retina_factor !== nothing
        rf = Array(retina_factor)  # (1, 3, n) as _spectrum_datastructs returns it
        for k in 1:n, c in 1:3
            w[3 * (k - 1) + c] = rf[1, c, k]
        end
    end
    LightMapParams(n, length(δ), n, WEIGHTS_IDENTITY, width, height, padded(spectrum, MAX_LAMBDA),
                   padded(δ, MAX_DELTA), Tuple(w), (0, 0, 0), ntuple(_ -> 0f0, 3 * MAX_CURVE),
                   ntuple(_ -> 0f0, 3 * MAX_CURVE))
end

struct Pinhole
    pos::NTuple{3,Float32}; dir::NTuple{3,Float32}; up::NTuple{3,Float32}; right::NTuple{3,Float32}
    fov_half_sin::Float32
    height::Int32; width::Int32; samples_per_pixel::Int32; order::Int32; lambda_mode::Int32
    lambda_min::Float32; lambda_max::Float32; jitter::Float32; _pad::Int32
    seed::UInt64
end
# camera_ray's camera (src/ray_generators.jl:18-50) at lambda 550 with dv * 0.025 / max(h, w)
Pinhole(cam::Cam, height, width; seed = 0, λ = 550f0) =
    Pinhole(Tuple(cam.pos), Tuple(cam.dir), Tuple(cam.up), Tuple(cam.right), cam.FOV_half_sin, height, width, 1,
            0, 0, λ, λ, 0.025f0, 0, seed)

struct LightRays
    upscale::Int32; lambda_mode::Int32; lambda_min::Float32; lambda_max::Float32
    jitter::Float32; _pad::Int32
    seed::UInt64
end
LightRays(upscale; lambda_mode = 0, lambda_min = 550f0, lambda_max = 550f0, jitter = 1f0, seed = 0) =
    LightRays(upscale, lambda_mode, lambda_min, lambda_max, jitter, 0, seed)

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
    counts, bv_tris = Main.pack_bv_tris(Array, Array(tris), bvs, memberships)  # host copies, once per scene
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
# place of bv_partition + pack_bv_tris on the host: ~0.5 ms for 10k triangles, so an animated
# scene (src/scenes/splash_scene.jl:41) can rebuild every frame with rebuild! (same workspace).
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

const HIPHitter = Union{HIPBruteForceHitter,HIPBVHitter}

# ---- the tracer (ExperimentalTracer's fields, src/alg_types.jl:101-115, plus the workspaces) ---
mutable struct HIPTracer <: AbstractTracer
    hit_idx::ROCArray{Int32}
    rndm::ROCArray{Float32}
    ray_swap::ROCArray{ADRay}
    δ
    # device workspaces, grown on first use and reused by every later call
    hit_ws::ROCArray{UInt8}
    part_ws::ROCArray{UInt8}
    run_ws::ROCArray{UInt8}
    reorder_ws::ROCArray{UInt8}
    d_count::ROCArray{Int64}   # live count kept on the device between bounces
    h_count::Vector{Int64}     # partition!'s host copy of it
end
HIPTracer(A, rays, N) = HIPTracer(ROCArray{Int32}(undef, length(rays)), ROCArray{Float32}(undef, length(rays)),
                                  similar(rays), upres(N), ROCArray{UInt8}(undef, 0), ROCArray{UInt8}(undef, 0),
                                  ROCArray{UInt8}(undef, 0), ROCArray{UInt8}(undef, 0), ROCArray{Int64}(undef, 1),
                                  Int64[0])

# the workspace `f` of at least `need` bytes (grown once, then reused)
function workspace!(tracer::HIPTracer, f::Symbol, need)
    ws = getfield(tracer, f)
    if length(ws) < need
        ws = ROCArray{UInt8}(undef, need)
        setfield!(tracer, f, ws)
    end
    ws
end
ws_bytes(sym, R) = ccall((sym, LIB), Csize_t, (Int64,), R)

# ---- operators ---------------------------------------------------------------------------------
# next_hit!(tracer, hitter, rays, n_tris) -- src/hitters.jl:24-30 semantics (ties to the lowest index)
function next_hit!(tracer::HIPTracer, hitter::HIPHitter, rays::ROCArray{ADRay}, n_tris)
    ws = workspace!(tracer, :hit_ws, ws_bytes(:splat_next_hit_ws_bytes, length(rays)))
    check(ccall((:splat_next_hit, LIB), Int32,
                (Ptr{Cvoid}, Int64, Ptr{Int64}, Ref{HitSet}, Ptr{Int32}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                pointer(rays), length(rays), C_NULL, hitter.hs, pointer(tracer.hit_idx), pointer(ws), length(ws),
                stream_ptr()), "next_hit")
end

# rays .= evolve_ray.(rays, tracer.hit_idx, tris[tracer.hit_idx], tracer.rndm, first_diffuse)
# (src/tracer.jl:181-201, the broadcast at :227 and :261): one bounce of the dual-number optics.
# force_rand as in run_evolution!; otherwise tracer.rndm (filled by the caller, as rand! does).
function evolve!(tracer::HIPTracer, rays::ROCArray{ADRay}, tris::ROCArray{FTri}; first_diffuse, force_rand = nothing,
                 optics = Optics(), seed = 0, bounce = 1)
    fr = isnothing(force_rand) ? NaN32 : Float32(force_rand)
    check(ccall((:splat_evolve, LIB), Int32,
                (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int32}, Ptr{Cvoid}, Int32, Ptr{Float32}, Float32, UInt64, Int32,
                 Int32, Ref{Optics}, Ptr{Cvoid}),
                pointer(rays), length(rays), C_NULL, pointer(tracer.hit_idx), pointer(tris), length(tris),
                pointer(tracer.rndm), fr, seed, bounce, first_diffuse, optics, stream_ptr()), "evolve")
end

# partition!(rays, swap; by = r.status != ACTIVE) -- src/partition.jl:122-132, stable: ACTIVE rays
# first in order, then the rest in order; returns the ACTIVE count (one host sync, as the
# reference's Array(...) copy does)
function partition!(tracer::HIPTracer, vals::ROCArray{ADRay}, swap::ROCArray{ADRay} = tracer.ray_swap)
    R = length(vals)
    ws = workspace!(tracer, :part_ws, ws_bytes(:splat_partition_ws_bytes, R))
    check(ccall((:splat_partition_active, LIB), Int32,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                pointer(vals), pointer(swap), R, pointer(tracer.d_count), tracer.h_count, pointer(ws), length(ws),
                stream_ptr()), "partition")
    return tracer.h_count[1]
end

# run_evolution!(hitter, ::ExperimentalTracer; ...) -- src/tracer.jl:233-282 fused on the device:
# the live count never leaves it (no host sync per bounce). Leaves rays and tracer.hit_idx as the
# reference's loop plus its realigning next_hit! does; reorder = true moves them to pixel order.
function run_evolution!(hitter::HIPHitter, tracer::HIPTracer; depth, first_diffuse, tris, rays, force_rand = nothing,
                        seed = 0, optics = Optics(), reorder = false, kwargs...)
    R = length(rays)
    ws = workspace!(tracer, :run_ws, ws_bytes(:splat_run_ws_bytes, R))
    p = RunParams(depth, first_diffuse, force_rand, seed, optics)
    check(ccall((:splat_run_evolution, LIB), Int32,
                (Ptr{Cvoid}, Int64, Ref{HitSet}, Ptr{Cvoid}, Int32, Ptr{Int32}, Ref{RunParams}, Ptr{Cvoid}, Csize_t,
                 Ptr{Int64}, Ptr{Cvoid}),
                pointer(rays), R, hitter.hs, pointer(tris), length(tris), pointer(tracer.hit_idx), p,
                pointer(ws), length(ws), pointer(tracer.d_count), stream_ptr()), "run_evolution")
    reorder && reorder!(tracer, rays)
    return
end

# run_evolution!(...; reorder = true)'s reorder (src/tracer.jl:270-278)
function reorder!(tracer::HIPTracer, rays::ROCArray{ADRay})
    R = length(rays)
    ws = workspace!(tracer, :reorder_ws, ws_bytes(:splat_reorder_ws_bytes, R))
    check(ccall((:splat_reorder_by_dest, LIB), Int32,
                (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                pointer(rays), R, pointer(tracer.hit_idx), pointer(ws), length(ws), stream_ptr()), "reorder")
end

# continuum_light_map2!(; tracer, tris, rays, spectrum, tex, ...) -- src/ray_imagers.jl:142-165:
# final_evolution of the DIFFUSE rays (mutates rays), then every splat (lambda_k, dx, dy) of
# atomic_spectrum_kernel2 (:78-113) accumulated into tex[i, j, k] (tex as tex_f() made it).
function continuum_light_map2!(; tracer::HIPTracer, tris::ROCArray{FTri}, rays::ROCArray{ADRay}, spectrum,
                               tex::ROCArray{Float32,3}, kwargs...)
    w, h, n = size(tex)
    n == length(spectrum) || throw(DimensionMismatch("tex has $n planes for $(length(spectrum)) wavelengths"))
    lm = LightMapParams(w, h, vec(Array(spectrum)), tracer.δ)
    check(ccall((:splat_light_map, LIB), Int32,
                (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int32}, Ptr{Cvoid}, Int32, Ref{LightMapParams}, Ptr{Float32},
                 Ptr{Cvoid}),
                pointer(rays), length(rays), C_NULL, pointer(tracer.hit_idx), pointer(tris), length(tris), lm,
                pointer(tex), stream_ptr()), "light_map")
end

# continuum_shade!(::ExperimentalImager2; tracer, RGB3, tris, rays, spectrum, retina_factor, tex,
# skydome) (src/ray_imagers.jl:440-503): rays in pixel order on a height-row grid; RGB3 is
# (n_pixels, 3) Float32; tex the light map as a plain array (bilinear reads, no texture object);
# skydome an RGB Float32 (w, h, 3) array or nothing
function continuum_shade!(::ExperimentalImager2; tracer::HIPTracer, RGB3::ROCArray{Float32}, tris::ROCArray{FTri},
                          rays::ROCArray{ADRay}, spectrum, retina_factor, tex::ROCArray{Float32,3}, skydome, height,
                          kwargs...)
    w, h, n = size(tex)
    lm = LightMapParams(w, h, vec(Array(spectrum)), tracer.δ; retina_factor)
    sky = skydome === nothing ? C_NULL : pointer(skydome)
    sw = skydome === nothing ? 0 : size(skydome, 1)
    sh = skydome === nothing ? 0 : size(skydome, 2)
    check(ccall((:splat_shade, LIB), Int32,
                (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Cvoid}, Int32, Ref{LightMapParams}, Ptr{Float32}, Ptr{Float32},
                 Int32, Int32, Int64, Ptr{Float32}, Ptr{Cvoid}),
                pointer(rays), length(rays), pointer(tracer.hit_idx), pointer(tris), length(tris), lm, pointer(tex),
                sky, sw, sh, height, pointer(RGB3), stream_ptr()), "shade")
end

# ---- ray sources -------------------------------------------------------------------------------
# wrap_ray_gen(camera_ray ...) (src/ray_generators.jl:18-62) into `rays` (height * width of them,
# column-major): the jitter dv from `seed` instead of CUDA.rand
function camera_rays!(rays::ROCArray{ADRay}, cam::Cam, height, width; seed = 0)
    length(rays) == height * width || throw(DimensionMismatch("rays"))
    check(ccall((:splat_raygen_pinhole, LIB), Int32, (Ref{Pinhole}, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                Pinhole(cam, height, width; seed), pointer(rays), length(rays), stream_ptr()), "raygen")
    rays
end

# rays_from_lights(lights, upscale) (src/ray_generators.jl:64-113) into `rays`; the Vector{RectLight}
# is passed as is (64-byte isbits records); rand(Float32)'s cell offsets come from `seed`
lights_ray_count(lights::Vector{RectLight}, upscale) =
    ccall((:splat_lights_ray_count, LIB), Int64, (Ptr{RectLight}, Int32, Int32), lights, length(lights), upscale)
function rays_from_lights!(rays::ROCArray{ADRay}, lights::Vector{RectLight}, upscale; kw...)
    p = LightRays(upscale; kw...)
    R = lights_ray_count(lights, upscale)
    R < 0 && throw(SplatError(Int32(R), "rays_from_lights"))
    length(rays) == R || throw(DimensionMismatch("rays"))
    check(ccall((:splat_raygen_lights, LIB), Int32,
                (Ptr{RectLight}, Int32, Ref{LightRays}, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                lights, length(lights), p, pointer(rays), R, stream_ptr()), "rays_from_lights")
    rays
end

# ---- trace! (src/tracer.jl:284-388) for HIPTracer ---------------------------------------------
# The reference's keyword interface; every array is a ROCArray and every step a SplatHIP method.
# Buffers (rays, tracers with their workspaces, RGB3) persist in a context keyed by the frame's
# sizes, so the frame loop of a scene script allocates only tex = tex_f() (as the reference does).
struct TraceContext
    fwd_rays::ROCArray{ADRay}
    fwd::HIPTracer
    bwd_rays::ROCArray{ADRay}
    bwd::HIPTracer
    RGB3::ROCArray{Float32,2}
end
const CONTEXTS = Dict{Tuple{Int,Int,Int,Int,Int},TraceContext}()
function trace_context(lights, forward_upscale, backward_upscale, height, width)
    Rf = lights_ray_count(lights, forward_upscale)
    Rf < 0 && throw(SplatError(Int32(Rf), "rays_from_lights"))
    h, w = height ÷ backward_upscale, width ÷ backward_upscale
    get!(CONTEXTS, (Rf, forward_upscale, h, w, backward_upscale)) do
        fr = ROCArray{ADRay}(undef, Rf)
        br = ROCArray{ADRay}(undef, h * w)
        TraceContext(fr, HIPTracer(ROCArray, fr, forward_upscale), br, HIPTracer(ROCArray, br, backward_upscale),
                     ROCArray{Float32}(undef, h * w * backward_upscale^2, 3))
    end
end

function trace!(::Type{HIPTracer}, ::Type{ExperimentalImager2}; forward_hitter::HIPHitter, backward_hitter::HIPHitter,
                cam, lights, forward_upscale, backward_upscale, tex_f, tris::ROCArray{FTri}, λ_min, dλ, λ_max,
                width, height, depth, first_diffuse, skydome, force_rand = nothing, iterations_per_frame = 1,
                seed = 0, optics = Optics(), kwargs...)
    ctx = trace_context(lights, forward_upscale, backward_upscale, height, width)
    spectrum, retina_factor = Main._spectrum_datastructs(Array, λ_min:dλ:λ_max)
    out = nothing
    for frame_iter in 1:iterations_per_frame
        s = UInt64(seed + 2 * (frame_iter - 1))
        tex = tex_f()  # ROCArray{Float32,3}: width x height x n_lambda
        # forward: light rays -> bounce loop -> light map (the reference's @async task, :317-344)
        rays_from_lights!(ctx.fwd_rays, lights, forward_upscale; seed = s)
        run_evolution!(forward_hitter, ctx.fwd; depth, first_diffuse, tris, rays = ctx.fwd_rays, force_rand,
                       seed = s, optics)
        continuum_light_map2!(; tracer = ctx.fwd, tris, rays = ctx.fwd_rays, spectrum, tex)
        # backward: camera rays -> bounce loop -> pixel order -> shading (:346-378), same stream
        camera_rays!(ctx.bwd_rays, cam, height ÷ backward_upscale, width ÷ backward_upscale; seed = s + 1)
        run_evolution!(backward_hitter, ctx.bwd; depth, first_diffuse, tris, rays = ctx.bwd_rays, force_rand,
                       seed = s + 1, optics, reorder = true)
        ctx.RGB3 .= 0
        continuum_shade!(ExperimentalImager2(); tracer = ctx.bwd, RGB3 = ctx.RGB3, tris, rays = ctx.bwd_rays,
                         spectrum, retina_factor, tex, skydome, height = height ÷ backward_upscale)
        AMDGPU.synchronize()
        # iterations_per_frame averaging (:380-384)
        if frame_iter == 1
            out = copy(ctx.RGB3)
        else
            out .= out .* (frame_iter - 1) / frame_iter .+ ctx.RGB3 ./ frame_iter
        end
    end
    return out
end

end # module
