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

# Model id of each GaussMLE PSF type (enum gmle_model), and its 8-float
# pixel-unit parameter block (the `psf` argument of gmle_fit): GaussianXYNB
# [σ, 0, ...], AstigmaticXYZNB [σx₀, σy₀, Ax, Ay, Bx, By, γ, d], zeros for the
# fitted-width models.  Pass the PSF after to_pixel_units
# (src/psf_models.jl:311-331): σ, σx₀, σy₀ in pixels, γ and d as given.
# The PSF types are GaussMLE.jl's (src/psf_models.jl:39, 64, 91, 141); this file
# is included inside module GaussMLE after psf_models.jl, so they are the
# parent module's.  Loaded on its own, only the fallbacks exist.
model_id(psf) = throw(ArgumentError("unsupported PSF model $(typeof(psf))"))
psf_block(psf) = throw(ArgumentError("unsupported PSF model $(typeof(psf))"))

const _GMLE = parentmodule(@__MODULE__)
if isdefined(_GMLE, :GaussianXYNB) && isdefined(_GMLE, :GaussianXYNBS) &&
   isdefined(_GMLE, :GaussianXYNBSXSY) && isdefined(_GMLE, :AstigmaticXYZNB)
    model_id(::_GMLE.GaussianXYNB) = Cint(0)
    model_id(::_GMLE.GaussianXYNBS) = Cint(1)
    model_id(::_GMLE.GaussianXYNBSXSY) = Cint(2)
    model_id(::_GMLE.AstigmaticXYZNB) = Cint(3)

    psf_block(p::_GMLE.GaussianXYNB) = Float32[p.σ, 0, 0, 0, 0, 0, 0, 0]
    psf_block(::_GMLE.GaussianXYNBS) = zeros(Float32, 8)
    psf_block(::_GMLE.GaussianXYNBSXSY) = zeros(Float32, 8)
    psf_block(p::_GMLE.AstigmaticXYZNB) = Float32[p.σx₀, p.σy₀, p.Ax, p.Ay, p.Bx, p.By, p.γ, p.d]
end

"""
    hip_fit!(results, uncertainties, log_likelihoods, data, psf_pixels, scmos,
             variance_map, x_corners, y_corners, constraints, iterations; kwargs...)

The reference kernel's own argument list (src/unified_kernel.jl:285-297) with
the PSF struct and `ParameterConstraints` as GaussMLE.jl holds them: the model
id and PSF block come from `model_id` / `psf_block`, the constraint vectors
from the struct's `lower` / `upper` / `max_step` fields (src/constraints.jl:6-10).
"""
function hip_fit!(results::Matrix{Float32}, uncertainties::Matrix{Float32}, log_likelihoods::Vector{Float32},
                  data::Array{Float32,3}, psf_pixels, scmos::Bool, variance_map::Matrix{Float32},
                  x_corners::Vector{Int32}, y_corners::Vector{Int32}, constraints, iterations::Integer; kwargs...)
    return hip_fit!(results, uncertainties, log_likelihoods, data, model_id(psf_pixels), psf_block(psf_pixels),
                    scmos, variance_map, x_corners, y_corners, Vector{Float32}(constraints.lower),
                    Vector{Float32}(constraints.upper), Vector{Float32}(constraints.max_step), iterations;
                    kwargs...)
end

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

function check(rc)
    if rc != GMLE_OK
        msg = last_error()
        rc in (-1, -2, -3) ? throw(ArgumentError(msg)) : error("gaussmle-hip: " * msg)
    end
    return nothing
end

"""
    hip_fit_batch!(params, crlb, llr, models, boxes, data, psf, lower, upper, max_step,
                   scmos, variance_map, x_corners, y_corners, iterations;
                   precision=:fp32, lanes=nothing, streams=4)

A mixed set of ROIs in one call: `models`/`boxes` are Vector{Int8}(n) (model
ids as `model_id`, box sizes 3..16); `data` holds the ROIs back to back, each
box*box Float32 column-major; `psf` is Matrix{Float32}(8, 4) and the
constraints Matrix{Float32}(6, 4), column m+1 for model id m; outputs are
Matrix{Float32}(6, n) (each ROI's P values, then NaN) and Vector{Float32}(n).
Replaces one `fit` call per (model, box size) on such a set.
"""
function hip_fit_batch!(params::Matrix{Float32}, crlb::Matrix{Float32}, llr::Vector{Float32},
                        models::Vector{Int8}, boxes::Vector{Int8}, data::Vector{Float32},
                        psf::Matrix{Float32}, lower::Matrix{Float32}, upper::Matrix{Float32},
                        max_step::Matrix{Float32}, scmos::Bool, variance_map::Matrix{Float32},
                        x_corners::Vector{Int32}, y_corners::Vector{Int32}, iterations::Integer;
                        precision::Symbol = :fp32, lanes::Union{Nothing,Vector{Int32}} = nothing,
                        streams::Integer = 4)
    n = length(models)
    ny, nx = size(variance_map)
    lp = lanes === nothing ? Ptr{Int32}(C_NULL) : pointer(lanes)
    rc = GC.@preserve params crlb llr models boxes data psf lower upper max_step variance_map x_corners y_corners lanes begin
        ccall((:gmle_fit_batch, libgmle), Cint,
              (Int64, Ptr{Int8}, Ptr{Int8}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
               Cint, Ptr{Float32}, Cint, Cint, Ptr{Int32}, Ptr{Int32}, Cint, Cint, Ptr{Int32}, Cint,
               Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float64}, Ptr{Float64}),
              n, models, boxes, data, psf, lower, upper, max_step, scmos ? 1 : 0, variance_map, ny, nx,
              x_corners, y_corners, iterations, PRECISION[precision], lp, streams, params, crlb, llr,
              Ptr{Float64}(C_NULL), Ptr{Float64}(C_NULL))
    end
    check(rc)
    return params, crlb, llr
end

# ---------------------------------------------------------------------------
# Serving (gmle_service_*): requests live in pinned memory from hip_host_alloc.
# ---------------------------------------------------------------------------
struct Request
    n_rois::Int64
    models::Ptr{Int8}
    boxes::Ptr{Int8}
    x_corners::Ptr{Int32}
    y_corners::Ptr{Int32}
    adu::Ptr{UInt16}
    params::Ptr{Float32}
    crlb::Ptr{Float32}
    llr::Ptr{Float32}
end

"Pinned, device-visible host memory as a Vector{T} (free with hip_host_free)."
function hip_host_alloc(::Type{T}, n::Integer) where {T}
    p = ccall((:gmle_host_alloc, libgmle), Ptr{Cvoid}, (Csize_t,), n * sizeof(T))
    p == C_NULL && error("gaussmle-hip: " * last_error())
    return unsafe_wrap(Vector{T}, Ptr{T}(p), n; own = false)
end
hip_host_free(v::Vector) = check(ccall((:gmle_host_free, libgmle), Cint, (Ptr{Cvoid},), pointer(v)))

function hip_service(psf::Matrix{Float32}, lower::Matrix{Float32}, upper::Matrix{Float32},
                     max_step::Matrix{Float32}, lanes::Vector{Int32}; device = 0, hbm_budget = 4 << 30,
                     slots = 4, precision::Symbol = :fp32, iterations = 20)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:gmle_service_create, libgmle), Cint,
                (Cint, Csize_t, Cint, Cint, Cint, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                 Ptr{Int32}, Ptr{Ptr{Cvoid}}),
                device, hbm_budget, slots, PRECISION[precision], iterations, psf, lower, upper, max_step, lanes, h))
    return h[]
end

hip_service_calibrate(s, offset::Matrix{Float32}, gain::Matrix{Float32}, var::Matrix{Float32}) =
    check(ccall((:gmle_service_calibrate, libgmle), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Cint, Cint),
                s, offset, gain, var, size(var, 1), size(var, 2)))

function hip_service_submit(s, r::Request)
    t = ccall((:gmle_service_submit, libgmle), Int64, (Ptr{Cvoid}, Ref{Request}), s, r)
    t < 0 && check(Cint(t))
    return t
end

function hip_service_wait(s, ticket::Integer; timeout = 60.0)
    rc = ccall((:gmle_service_wait, libgmle), Cint, (Ptr{Cvoid}, Int64, Cdouble), s, ticket, timeout)
    rc == 1 && error("gaussmle-hip: request $ticket not done after $(timeout)s")
    return check(rc)
end

hip_service_destroy(s) = check(ccall((:gmle_service_destroy, libgmle), Cint, (Ptr{Cvoid},), s))

"""
    hip_getboxes(imagestack; boxsize = 7, overlap = 2.0, sigma_small = 1.0, sigma_large = 2.0, minval = 0.0)

Candidate boxes of a (rows, cols, frames) Float32 stack, as SMLMBoxer.getboxes
cuts them (examples/full_pipeline.jl:73-81): returns (rois, x_corners,
y_corners, frames) with 1-based corners, ordered by frame, then column-major
pixel index.
"""
function hip_getboxes(imagestack::Array{Float32,3}; boxsize::Integer = 7, overlap::Real = 2.0,
                      sigma_small::Real = 1.0, sigma_large::Real = 2.0, minval::Real = 0.0)
    ny, nx, nf = size(imagestack)
    r = max(1, ceil(Int, Float32(boxsize) - Float32(overlap)) - 1) + 1
    cap = nf * cld(ny, r) * cld(nx, r)   # at most one detection per r x r cell
    rois = Array{Float32,3}(undef, boxsize, boxsize, cap)
    xc = Vector{Int32}(undef, cap); yc = Vector{Int32}(undef, cap); fr = Vector{Int32}(undef, cap)
    n = Ref{Int64}(0)
    check(ccall((:gmle_find_boxes, libgmle), Cint,
                (Ptr{Cfloat}, Cint, Cint, Cint, Cint, Cfloat, Cfloat, Cfloat, Cfloat, Int64, Ptr{Cfloat},
                 Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Cfloat}, Ref{Int64}),
                imagestack, ny, nx, nf, boxsize, overlap, sigma_small, sigma_large, minval, cap, rois,
                xc, yc, fr, C_NULL, n))
    m = n[]
    return rois[:, :, 1:m], xc[1:m], yc[1:m], fr[1:m]
end

"""
    hip_chi2_pvalue(chi2, df)

`1 - Float32(cdf(Chisq(df), chi2))` as the device epilogue computes it
(src/interface.jl:265-275), for callers that finish fits on the host.
"""
hip_chi2_pvalue(chi2::Real, df::Integer) = ccall((:gmle_chi2_pvalue, libgmle), Cfloat, (Cfloat, Cint), chi2, df)

"""
    hip_epilogue!(out, params, crlb, llr, model, box, stride, x_corners, y_corners, psx, psy, n; stream = C_NULL)

Device arrays (`CuArray`-like pointers from a HIP array package, or raw
`Ptr`s): `out` is (10, n) Float32 with x, y, σx, σy in µm, the p-value,
fitted widths and their uncertainties in µm, and χ² per ROI
(src/roi_batch.jl:29-34, :125-302).  `params` / `crlb` are (stride, n);
C_NULL corners mean fit(Array)'s x = 1 + (k-1) box, y = 1.
"""
function hip_epilogue!(out, params, crlb, llr, model::Integer, box::Integer, stride::Integer, x_corners, y_corners,
                       psx::Real, psy::Real, n::Integer; stream = C_NULL)
    check(ccall((:gmle_epilogue_device, libgmle), Cint,
                (Int64, Cint, Cint, Ptr{Int8}, Ptr{Int8}, Cint, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat},
                 Ptr{Int32}, Ptr{Int32}, Cdouble, Cdouble, Ptr{Cfloat}, Ptr{Cvoid}),
                n, model, box, C_NULL, C_NULL, stride, params, crlb, llr, x_corners, y_corners, psx, psy, out,
                stream))
    return out
end

end # module
