# ParticleHolographyHIP.jl -- the reference-side binding: ParticleHolography.jl's holofunc.jl API
# (src/holofunc.jl:5-6 exports, src/types.jl:3) on AMDGPU.jl ROCArrays, over libholo's C ABI
# (include/holo.h) with `ccall`. Drop-in: same names, argument order and return types; the
# structs keep their `.data` field and 1-argument constructors (runtests.jl:10-60) and gain
# optional provenance (z, wavelength, pitch) that lets libholo generate H in FP64.
module ParticleHolographyHIP

using AMDGPU
using FixedPointNumbers
using Random
using UUIDs

export CuTransferSqrtPart, CuTransfer, CuWavefront, CuLowPassFilter
export cu_rectangle_filter, cu_super_gaussian_filter, cu_apply_low_pass_filter, cu_apply_low_pass_filter!
export cu_make_background_mode, cu_ingest_u8
export cu_connected_component_labeling, cu_find_valid_labels, slice_bounding_rectangles, cu_dilate,
       link_rectangles, particle_bounding_boxes, particle_bounding_boxes_3d
export cu_transfer_sqrt_arr, cu_transfer, cu_gabor_wavefront, cu_phase_retrieval_holo,
       cu_get_reconst_vol, cu_get_reconst_xyprojection, cu_get_reconst_vol_and_xyprojection,
       cu_get_reconst_complex_vol, cu_asm_prop!, cu_2d_pad, cu_get_reconst_vol_and_xyprojection_padded

const libholo = get(ENV, "HOLO_LIB", joinpath(@__DIR__, "..", "lib", "libholo.so"))

# ---- types.jl:13-61 ------------------------------------------------------------------------
struct CuTransferSqrtPart{T<:AbstractFloat}
    data::ROCArray{T,2}
    wavlen::Union{Nothing,Float64}
    dx::Union{Nothing,Float64}
end
CuTransferSqrtPart(data::ROCArray{T,2}) where {T} = CuTransferSqrtPart{T}(data, nothing, nothing)

struct CuTransfer{T<:Complex}
    data::ROCArray{T,2}
    z::Union{Nothing,Float64}
    wavlen::Union{Nothing,Float64}
    dx::Union{Nothing,Float64}
end
CuTransfer(data::ROCArray{T,2}) where {T} = CuTransfer{T}(data, nothing, nothing, nothing)

struct CuWavefront{T<:Complex}
    data::ROCArray{T,2}
end

struct CuLowPassFilter{T<:AbstractFloat}
    data::ROCArray{T,2}
end

for T in (CuTransferSqrtPart, CuTransfer, CuWavefront, CuLowPassFilter)
    @eval begin
        Base.size(x::$T) = size(x.data)
        Base.axes(x::$T) = axes(x.data)
        Base.eltype(x::$T) = eltype(x.data)
        Base.ndims(x::$T) = ndims(x.data)
    end
end

# ---- C ABI plumbing ----------------------------------------------------------------------------
struct HoloTF            # holo_tf
    data::Ptr{Cvoid}
    z::Float64
    wavlen::Float64
    dx::Float64
    generated::Cint
end

const HOLO_N0F8, HOLO_F16, HOLO_F32, HOLO_F64, HOLO_C64 = 0, 1, 2, 3, 4
const _handles = Dict{Int,Ptr{Cvoid}}()

function _handle()
    dev = AMDGPU.device_id(AMDGPU.device()) - 1
    get!(_handles, dev) do
        r = Ref{Ptr{Cvoid}}(C_NULL)
        _check(ccall((:holo_create, libholo), Cint, (Ptr{Ptr{Cvoid}}, Cint), r, dev))
        r[]
    end
end

_stream() = AMDGPU.stream().stream   # hipStream_t of the task-local stream

function _check(code::Cint)
    code == 0 && return nothing
    msg = unsafe_string(ccall((:holo_last_error, libholo), Cstring, ()))
    code == 1 && throw(AssertionError(msg))
    code == 3 && throw(ArgumentError(msg))
    error("libholo error $code: $msg")
end

_generated(t::CuTransfer) = t.z !== nothing && t.wavlen !== nothing && t.dx !== nothing
_tf(t::CuTransfer, gen::Bool) = gen ? HoloTF(C_NULL, t.z, t.wavlen, t.dx, 1) :
                                      HoloTF(Ptr{Cvoid}(pointer(t.data)), 0.0, 0.0, 0.0, 0)
_dtype(::Type{N0f8}) = HOLO_N0F8
_dtype(::Type{Float16}) = HOLO_F16
_dtype(::Type{Float32}) = HOLO_F32
_dtype(::Type{Float64}) = HOLO_F64
_dtype(T) = throw(ArgumentError("return_type must be Subtype of AbstractFloat or N0f8. Got $T"))

# ---- holofunc.jl -------------------------------------------------------------------------------
function cu_transfer_sqrt_arr(datlen::Int, wavlen::AbstractFloat, dx::AbstractFloat)
    plane = ROCArray{Float32}(undef, datlen, datlen)
    _check(ccall((:holo_transfer_sqrt, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cint, Ptr{Cvoid}),
                 _handle(), _stream(), datlen, datlen, wavlen, dx, HOLO_F32, pointer(plane)))
    return CuTransferSqrtPart{Float32}(plane, Float64(wavlen), Float64(dx))
end

function cu_transfer(z0::AbstractFloat, datLen::Int, wavLen::AbstractFloat, d_sqr::CuTransferSqrtPart)
    plane = ROCArray{ComplexF32}(undef, datLen, datLen)
    prov = d_sqr.dx !== nothing && d_sqr.wavlen == Float64(wavLen)
    mode = prov ? 2 : 1                       # HOLO_TF_FP64 / HOLO_TF_FP64_D
    dtype = eltype(d_sqr.data) == Float32 ? HOLO_F32 : HOLO_F64
    _check(ccall((:holo_transfer, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Ptr{Cvoid}, Cint, Cint, Cdouble, Ptr{Cvoid}),
                 _handle(), _stream(), datLen, datLen, z0, wavLen, pointer(d_sqr.data), dtype, mode,
                 prov ? d_sqr.dx : 0.0, pointer(plane)))
    return prov ? CuTransfer{ComplexF32}(plane, Float64(z0), Float64(wavLen), d_sqr.dx) : CuTransfer(plane)
end

function cu_gabor_wavefront(holo::ROCArray{Float32,2})
    out = ROCArray{ComplexF32}(undef, size(holo)...)
    _check(ccall((:holo_gabor_wavefront, libholo), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                 _handle(), _stream(), size(holo, 1), size(holo, 2), pointer(holo), pointer(out)))
    return CuWavefront(out)
end

function cu_phase_retrieval_holo(holo1::ROCArray{Float32,2}, holo2::ROCArray{Float32,2}, transfer::CuTransfer{ComplexF32},
                                 invtransfer::CuTransfer{ComplexF32}, priter::Int, datlen::Int)
    @assert size(holo1) == size(holo2) == size(transfer.data) == size(invtransfer.data) == (datlen, datlen) "All arrays must have the same size as ($datlen, $datlen). Got $(size(holo1)), $(size(holo2)), $(size(transfer.data)), $(size(invtransfer.data))."
    gen = _generated(transfer) && _generated(invtransfer)
    out = ROCArray{ComplexF32}(undef, datlen, datlen)
    _check(ccall((:holo_phase_retrieval, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ref{HoloTF}, Ref{HoloTF}, Cint, Ptr{Cvoid}, Cint),
                 _handle(), _stream(), datlen, datlen, pointer(holo1), pointer(holo2),
                 _tf(transfer, gen), _tf(invtransfer, gen), priter, pointer(out), 1))
    return CuWavefront(out)
end

function cu_2d_pad(inarr)
    nx, ny = size(inarr)
    out = ROCArray{ComplexF32}(undef, 2nx, 2ny)
    src = ROCArray{ComplexF32}(inarr)
    _check(ccall((:holo_2d_pad, libholo), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                 _handle(), _stream(), nx, ny, pointer(src), pointer(out)))
    return out
end

function _sizes_ok(wavefront, tf_front, tf_dz)
    @assert size(wavefront.data) == size(tf_front.data) == size(tf_dz.data) "All arrays must have the same size. Got $(size(wavefront.data)), $(size(tf_front.data)), $(size(tf_dz.data))."
end

function cu_get_reconst_xyprojection(wavefront::CuWavefront{ComplexF32}, transfer_front::CuTransfer{ComplexF32},
                                     transfer_dz::CuTransfer{ComplexF32}, slices::Int)
    _sizes_ok(wavefront, transfer_front, transfer_dz)
    gen = _generated(transfer_front) && _generated(transfer_dz)
    nx, ny = size(wavefront.data)
    proj = ROCArray{Float32}(undef, nx, ny)
    _check(ccall((:holo_reconst_xyprojection, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Ref{HoloTF}, Ref{HoloTF}, Cint, Ptr{Cvoid}),
                 _handle(), _stream(), nx, ny, pointer(wavefront.data), _tf(transfer_front, gen),
                 _tf(transfer_dz, gen), slices, pointer(proj)))
    return proj
end

function _reconst_vol(wavefront, transfer_front, transfer_dz, slices, T, code, with_proj)
    _sizes_ok(wavefront, transfer_front, transfer_dz)
    gen = _generated(transfer_front) && _generated(transfer_dz)
    nx, ny = size(wavefront.data)
    vol = ROCArray{T}(undef, nx, ny, slices)
    proj = with_proj ? ROCArray{Float32}(undef, nx, ny) : nothing
    _check(ccall((:holo_reconst_vol, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Ref{HoloTF}, Ref{HoloTF}, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                 _handle(), _stream(), nx, ny, pointer(wavefront.data), _tf(transfer_front, gen), _tf(transfer_dz, gen),
                 slices, code, pointer(vol), with_proj ? pointer(proj) : C_NULL))
    return vol, proj
end

cu_get_reconst_vol(wavefront::CuWavefront{ComplexF32}, transfer_front::CuTransfer{ComplexF32},
                   transfer_dz::CuTransfer{ComplexF32}, slices::Int, return_type::Type=N0f8) =
    _reconst_vol(wavefront, transfer_front, transfer_dz, slices, return_type, _dtype(return_type), false)[1]

cu_get_reconst_complex_vol(wavefront::CuWavefront{ComplexF32}, transfer_front::CuTransfer{ComplexF32},
                           transfer_dz::CuTransfer{ComplexF32}, slices::Int) =
    _reconst_vol(wavefront, transfer_front, transfer_dz, slices, ComplexF32, HOLO_C64, false)[1]

cu_get_reconst_vol_and_xyprojection(wavefront::CuWavefront{ComplexF32}, transfer_front::CuTransfer{ComplexF32},
                                    transfer_dz::CuTransfer{ComplexF32}, slices::Int, return_type::Type=N0f8) =
    _reconst_vol(wavefront, transfer_front, transfer_dz, slices, return_type, _dtype(return_type), true)

function cu_get_reconst_vol_and_xyprojection_padded(wavefront::CuWavefront{ComplexF32}, transfer_front::CuTransfer{ComplexF32},
                                                    transfer_dz::CuTransfer{ComplexF32}, slices::Int, return_type::Type=N0f8)
    expected_size = map(x -> 2 * x, size(wavefront.data))
    @assert expected_size == size(transfer_front.data) == size(transfer_dz.data) "size(transfer_front.data) and size(transfer_dz.data) must be equal to 2*size(wavefront.data). Got $(size(wavefront.data)), $(size(transfer_front.data)), $(size(transfer_dz.data))."
    _dtype(return_type)
    gen = _generated(transfer_front) && _generated(transfer_dz)
    nx, ny = size(wavefront.data)
    vol = ROCArray{N0f8}(undef, nx, ny, slices)
    proj = ROCArray{Float32}(undef, nx, ny)
    _check(ccall((:holo_reconst_vol_and_xyprojection_padded, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Ref{HoloTF}, Ref{HoloTF}, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                 _handle(), _stream(), nx, ny, pointer(wavefront.data), _tf(transfer_front, gen), _tf(transfer_dz, gen),
                 slices, pointer(vol), pointer(proj)))
    return vol, proj
end

function cu_asm_prop!(outholo::CuWavefront, inholo::CuWavefront, d_sqr::CuTransferSqrtPart,
                      zprop::Float64, datlen::Int, λ::Float64)
    tf = d_sqr.dx !== nothing && d_sqr.wavlen == λ ?
         HoloTF(C_NULL, zprop, λ, d_sqr.dx, 1) :
         _tf(cu_transfer(zprop, datlen, λ, d_sqr), false)
    _check(ccall((:holo_asm_prop, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ref{HoloTF}),
                 _handle(), _stream(), datlen, datlen, pointer(outholo.data), pointer(inholo.data), tf))
    return nothing
end

# ---- frequency_filters.jl:31-107 -------------------------------------------------------------
const HOLO_LPF_RECTANGLE, HOLO_LPF_SUPER_GAUSSIAN = 0, 1

function _lpf(kind, prop_dist, wavlen, imglen, pixel_picth)
    arr = ROCArray{Float32}(undef, imglen, imglen)
    _check(ccall((:holo_lowpass_filter, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Float64, Float64, Cint, Float64, Ptr{Cvoid}),
                 _handle(), _stream(), kind, prop_dist, wavlen, imglen, pixel_picth, pointer(arr)))
    return CuLowPassFilter(arr)
end
cu_rectangle_filter(prop_dist::AbstractFloat, wavlen::AbstractFloat, imglen::Int, pixel_picth::AbstractFloat) =
    _lpf(HOLO_LPF_RECTANGLE, prop_dist, wavlen, imglen, pixel_picth)
cu_super_gaussian_filter(prop_dist::AbstractFloat, wavlen::AbstractFloat, imglen::Int, pixel_picth::AbstractFloat) =
    _lpf(HOLO_LPF_SUPER_GAUSSIAN, prop_dist, wavlen, imglen, pixel_picth)

function _apply_lpf!(out::ROCArray, holo::CuWavefront, lpf::CuLowPassFilter{Float32})
    @assert size(holo.data) == size(lpf.data) "low-pass filter and wavefront must have the same size"
    nx, ny = size(holo.data)
    _check(ccall((:holo_apply_low_pass_filter, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 _handle(), _stream(), nx, ny, pointer(out), pointer(holo.data), pointer(lpf.data)))
    return out
end
function cu_apply_low_pass_filter!(holo::CuWavefront, lpf::CuLowPassFilter)
    _apply_lpf!(holo.data, holo, lpf)
    return nothing
end
cu_apply_low_pass_filter(holo::CuWavefront, lpf::CuLowPassFilter) =
    CuWavefront(_apply_lpf!(similar(holo.data), holo, lpf))

# ---- utils.jl:286-317 (cu_make_background_mode) and 8-bit ingest ------------------------------
const HOLO_BACKGROUND_MODE, HOLO_BACKGROUND_MEAN = 0, 1

function cu_make_background_mode(grayimglist)
    H, W = size(grayimglist[1])
    stack = Array{UInt8}(undef, H, W, length(grayimglist))
    for (k, img) in enumerate(grayimglist)
        stack[:, :, k] .= eltype(img) === UInt8 ? img : round.(UInt8, clamp.(Float32.(img) .* 255, 0, 255))
    end
    d_stack = ROCArray(stack)
    out = ROCArray{Float64}(undef, H, W)
    _check(ccall((:holo_make_background, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                 _handle(), _stream(), H, W, length(grayimglist), pointer(d_stack), HOLO_BACKGROUND_MODE,
                 pointer(out)))
    return Array(out)
end

# Float32 holograms from raw 8-bit images on the device: v/255, or Float32((v/255 - bkg) + offset)
# with a Float64 background (preprocessings.md:21-23). No counterpart function in the reference.
function cu_ingest_u8(imgs::ROCArray{UInt8}, bkg::Union{Nothing,ROCArray{Float64,2}}=nothing,
                      offset::Float64=(bkg === nothing ? 0.0 : 64 / 255))
    nx, ny = size(imgs, 1), size(imgs, 2)
    count = ndims(imgs) == 2 ? 1 : size(imgs, 3)
    out = ROCArray{Float32}(undef, size(imgs))
    _check(ccall((:holo_ingest_u8, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Cvoid}),
                 _handle(), _stream(), nx, ny, count, pointer(imgs),
                 bkg === nothing ? C_NULL : pointer(bkg), offset, pointer(out)))
    return out
end

# ---- ccl.jl:186-281 and the per-slice part of particle_detection.jl:58-96 ----------------------
const HOLO_CCL_F32, HOLO_CCL_BINARY, HOLO_CCL_N0F8_LE = 0, 1, 2

struct HoloBox            # holo_box
    label::UInt32
    i0::Int32
    j0::Int32
    i1::Int32
    j1::Int32
end

_ccl_kind(::Type{Float32}) = HOLO_CCL_F32
_ccl_kind(::Type{<:Union{Bool,UInt8,N0f8}}) = HOLO_CCL_BINARY

# Labels (1 + the smallest linear index of the component, background 0), as the reference's
# union-find leaves them.
function cu_connected_component_labeling(input_img::ROCArray{T,2}) where {T}
    @assert length(input_img) <= 2^32 - 1 "Image is too large. Maximum length is 2^32-1."
    nx, ny = size(input_img)
    out = ROCArray{UInt32}(undef, nx, ny)
    _check(ccall((:holo_ccl_label, libholo), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}, Cint, Float64, Ptr{Cvoid}),
                 _handle(), _stream(), nx, ny, 1, pointer(input_img), _ccl_kind(T), 0.0, pointer(out)))
    return out
end

function _label_boxes(labels::ROCArray{UInt32,2}; max_boxes::Int=1 << 16)
    nx, ny = size(labels)
    while true
        boxes = ROCArray{HoloBox}(undef, max_boxes)
        nbox = ROCArray{Cint}(undef, 1)
        _check(ccall((:holo_label_boxes, libholo), Cint,
                     (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                     _handle(), _stream(), nx, ny, 1, pointer(labels), max_boxes, pointer(boxes), pointer(nbox)))
        n = Int(Array(nbox)[1])
        n <= max_boxes && return sort(Array(boxes)[1:n]; by=b -> b.label)
        max_boxes = n
    end
end

cu_find_valid_labels(labels::ROCArray{UInt32,2}) = Int64[b.label for b in _label_boxes(labels)]

# Rectangles (x_min, y_min, x_max, y_max) of the components of every slice, ascending label
# order per slice: what particle_bounding_boxes gets from cu_connected_component_labeling +
# cu_find_valid_labels + get_bounding_rectangles slice by slice, in one call without label
# images on the host. `threshold` binarises an N0f8 volume as d_vol .<= threshold.
function slice_bounding_rectangles(vol::ROCArray{T,3}; threshold=nothing, max_boxes::Int=4096) where {T}
    nx, ny, slices = size(vol)
    kind = threshold === nothing ? _ccl_kind(T) : HOLO_CCL_N0F8_LE
    thr = threshold === nothing ? 0.0 : Float64(threshold)
    while true
        boxes = ROCArray{HoloBox}(undef, max_boxes * slices)
        nbox = ROCArray{Cint}(undef, slices)
        _check(ccall((:holo_ccl_boxes, libholo), Cint,
                     (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}, Cint, Float64, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                     _handle(), _stream(), nx, ny, slices, pointer(vol), kind, thr, max_boxes,
                     pointer(boxes), pointer(nbox)))
        n = Array(nbox)
        if slices == 0 || maximum(n) <= max_boxes
            b = Array(boxes)
            return [[(Int(x.j0) + 1, Int(x.i0) + 1, Int(x.j1) + 1, Int(x.i1) + 1)
                     for x in sort(b[(z - 1) * max_boxes + 1:(z - 1) * max_boxes + n[z]]; by=x -> x.label)]
                    for z in 1:slices]
        end
        max_boxes = maximum(n)
    end
end

# gen_particle_neighborhoods + update_particle_neighborhoods(3d)! per slice +
# finalize_particle_neighborhoods! (ccl.jl:266-466) in libholo's native linking; keys drawn as
# the reference draws them, uuid1 of a fresh MersenneTwister(1234) per start slice.
function link_rectangles(rects::Vector{<:AbstractVector}; three_d::Bool=false)
    nslices = length(rects)
    nslices == 0 && return Dict{UUID,Vector{Int}}()
    offsets = zeros(Int64, nslices + 1)
    for s in 1:nslices
        offsets[s+1] = offsets[s] + length(rects[s])
    end
    total = offsets[end]
    flat = Vector{Int32}(undef, 4 * total)
    k = 0
    for r in rects, br in r, c in 1:4
        k += 1
        flat[k] = br[c]
    end
    cap = max(total, 1)
    rows = Matrix{Int64}(undef, 6, cap)
    born = Vector{Int32}(undef, cap)
    alive = Vector{UInt8}(undef, cap)
    nrows = Ref{Int64}(0)
    _check(ccall((:holo_link_rectangles, libholo), Cint,
                 (Ptr{Int32}, Ptr{Int64}, Cint, Cint, Ptr{Int64}, Ptr{Int32}, Ptr{UInt8}, Ptr{Int64}),
                 flat, offsets, nslices, three_d ? 1 : 0, rows, born, alive, nrows))
    out = Dict{UUID,Vector{Int}}()
    rng, cur = MersenneTwister(1234), Int32(0)
    for i in 1:nrows[]
        if born[i] != cur
            rng, cur = MersenneTwister(1234), born[i]
        end
        key = uuid1(rng)
        alive[i] == 1 && (out[key] = Vector{Int}(rows[:, i]))
    end
    return out
end

# particle_detection.jl:58-96 with the per-slice rectangles from one libholo call and the
# native linking; `threshold` fuses d_vol .<= threshold for an N0f8 volume.
particle_bounding_boxes(d_bin_vol::ROCArray{T,3}; threshold=nothing) where {T} =
    link_rectangles(slice_bounding_rectangles(d_bin_vol; threshold=threshold); three_d=false)
particle_bounding_boxes_3d(d_bin_vol::ROCArray{T,3}; threshold=nothing) where {T} =
    link_rectangles(slice_bounding_rectangles(d_bin_vol; threshold=threshold); three_d=true)

function cu_dilate(vol::ROCArray{Bool,3}; blocksize=32)
    nx, ny, slices = size(vol)
    out = similar(vol)
    _check(ccall((:holo_dilate, libholo), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                 _handle(), _stream(), nx, ny, slices, pointer(vol), pointer(out)))
    return out
end

end # module
