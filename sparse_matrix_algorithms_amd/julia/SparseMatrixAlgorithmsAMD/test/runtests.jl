# Lap1D known answers on one MI355X (the reference's own harness cases,
# polfed.jl:1220-1233 and chebyshev_subspace_iteration.jl:903, 980-992).
# Needs libsmalgo.so built (python -c "import __graft_entry__ as g; g.build()")
# or SMALGO_LIB pointing at it.
using Test, SparseArrays, LinearAlgebra
using SparseMatrixAlgorithmsAMD

lap1d(n) = spdiagm(-1 => fill(-1.0, n - 1), 0 => fill(2.0, n), 1 => fill(-1.0, n - 1))
lap1d_eigs(n) = [2 - 2cos(k * pi / (n + 1)) for k in 1:n]

@testset "POLFED Lap1D, delta window at the centre" begin
    n, nev, s, K = 10_000, 1000, 64, 32
    λall = lap1d_eigs(n)
    i1 = clamp(round(Int, n / 2) - nev ÷ 2, 1, n)
    i2 = clamp(i1 + nev - 1, 1, n)
    emin, emax = λall[i1], λall[i2]
    λ, X, res = polfed(lap1d(n), nev, K; s=s, smin=0.0, smax=4.0, sigma=(emin + emax) / 2,
                       emin=emin, emax=emax, mmax=79, window_type=:delta, show_progress=false)
    @test length(λ) == nev
    @test maximum(res) < 1e-10
    @test maximum(abs.(sort(λ) .- λall[i1:i2])) < 1e-10
end

@testset "Chebyshev Lap1D, matrix-free form" begin
    n = 2000
    A = lap1d(n)
    λall = lap1d_eigs(n)
    matvec!(y, x) = mul!(y, A, x)
    λ, X, res = chebyshev_poly_real_symm_sparse(matvec!, n, λall[981], λall[1020];
                                               matblock=(Y, X) -> mul!(Y, A, X), smin=0.0,
                                               smax=4.0, m0=80, degree=128, maxiter=10,
                                               tol=1e-8, window_type=:delta,
                                               σ=(λall[981] + λall[1020]) / 2)
    @test all(r -> r < 1e-6, res)
    @test length(λ) >= 38
end
