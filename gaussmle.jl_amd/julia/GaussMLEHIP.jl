"""
    GaussMLEHIP

Julia binding of libgaussmle_hip.so (include/gaussmle.h) for GaussMLE.jl.

`hip_fit!` takes exactly the arrays GaussMLE.jl passes to
`unified_gaussian_mle_kernel!` (src/unified_kernel.jl:285-297) and fills the
same outputs, so the GPU branches of `fit` in src/interface.jl
(:217-263, :334-370, :421-457) reduce to one call each.  `hip_functional()`
replaces the `CUDA.functional()` probe of src/devices.jl:11-35.

Not runnable in the build container (no Julia toolchain there); the C ABI it
binds is exercised by the repository's Python tests.
"""
module GaussMLEHIP

const libgmle = get(ENV, "GAUSSMLE_HIP_LIB", joinpath(@__DIR__, "..", "lib", "libgaussmle_hip.so"))

const GMLE_OK = Cint(0)
const PRECISION = Dict(:fp32 => Cint(0), :mixed => Cint(1))

hip_device_count() = ccall((:gmle_device_count, libgmle), Cint, ())
hip_functional() = hip_device_count() > 0

function last_error()
    buf = Vector{UInt8}(undef, 1024)
    ccall((:gmle_last_error, libgmle), Csize_t, (Ptr{UInt8}, Csize_t), buf, length(buf))
    return unsafe_string(pointer(buf))
end

# Model id of each GaussMLE PSF type, and its 8-float pixel-unit parameter block.
model_id(::Any) = error("unsupported PSF model")
psf_block(psf) = zeros(Float32, 8)

"""
    hip_fit!(results, uncertainties, log_likelihoods, data, model, psf_block,
             scmos::Bool, variance_map, x_corners, y_corners,
             lower, upper, max_step, iterations; precision=:fp32, lanes=0,
             crlb64=nothing, llr64=nothing)

`results`/`uncertainties`: Matrix{Float32}(P, n); `log_likelihoods`:
Vector{Float32}(n); `data`: Array{Float32,3}(box, box, n); `variance_map`:
Matrix{Float32}(ny, nx) indexed [row, col]; corners: Vector{Int32}, 1-based.
Throws `ArgumentError` for rejected input, `ErrorException` otherwise.
"""
function hip_fit!(results::Matrix{Float32}, uncertainties::Matrix{Float32}, log_likelihoods::Vector{Float32},
                  data::Array{Float32,3}, model::Integer, psf::Vector{Float32}, scmos::Bool,
                  variance_map::Matrix{Float32}, x_corners::Vector{Int32}, y_corners::Vector{Int32},
                  lower::Vector{Float32}, upper::Vector{Float32}, max_step::Vector{Float32}, iterations::Integer;
                  precision::Symbol = :fp32, lanes::Integer = 0,
                  crlb64::Union{Nothing,Matrix{Float64}} = nothing,
                  llr64::Union{Nothing,Vector{Float64}} = nothing)
    box = size(data, 1)
    n = size(data, 3)
    ny, nx = size(variance_map)
    c64 = crlb64 === nothing ? Ptr{Float64}(C_NULL) : pointer(crlb64)
    l64 = llr64 === nothing ? Ptr{Float64}(C_NULL) : pointer(llr64)
    rc = GC.@preserve results uncertainties log_likelihoods data psf variance_map x_corners y_corners lower upper max_step crlb64 llr64 begin
        ccall((:gmle_fit, libgmle), Cint,
              (Cint, Ptr{Float32}, Cint, Int64, Ptr{Float32}, Cint, Ptr{Float32}, Cint, Cint,
               Ptr{Int32}, Ptr{Int32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Cint, Cint, Cint,
               Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float64}, Ptr{Float64}),
              model, psf, box, n, data, scmos ? 1 : 0, variance_map, ny, nx,
              x_corners, y_corners, lower, upper, max_step, iterations, PRECISION[precision], lanes,
              results, uncertainties, log_likelihoods, c64, l64)
    end
    if rc != GMLE_OK
        msg = last_error()
        rc in (-1, -2, -3) ? throw(ArgumentError(msg)) : error("gaussmle-hip: " * msg)
    end
    return results, uncertainties, log_likelihoods
end

end # module
