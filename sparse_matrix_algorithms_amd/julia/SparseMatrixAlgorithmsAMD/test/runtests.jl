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

@testset "coefficients and apply_poly_block! against the spectral form" begin
    @test jackson_kernel(0) == [1.0]
    g = jackson_kernel(40)
    @test length(g) == 41 && g[1] ≈ 1.0
    # a Chebyshev series evaluated at cos θ from its coefficients
    cheb(c, x) = sum(c[k+1] * cos(k * acos(clamp(x, -1, 1))) for k in 0:length(c)-1)
    c = cheb_window_coeffs(-0.3, 0.2, 200)
    @test cheb(c, 0.0) ≈ 1.0 atol = 0.02
    @test abs(cheb(c, 0.8)) < 0.02
    n = 300
    A = lap1d(n)
    λall = lap1d_eigs(n)
    V = [sqrt(2 / (n + 1)) * sin(i * k * pi / (n + 1)) for i in 1:n, k in 1:n]   # eigenvectors
    co = SparseMatrixAlgorithmsAMD._build_coeffs(60, 0.0, 4.0, 2.0, 1.8, 2.2; window_type=:step)
    Q = randn(n, 5)
    Y = similar(Q)
    apply_poly_block!(Y, A, Q, co, 0.0, 4.0, POLFEDWorkspace(n, 5))
    p = [cheb(co, (λ - 2.0) / 2.0) for λ in λall]
    @test maximum(abs.(Y .- V * (p .* (V' * Q)))) < 1e-12 * maximum(abs.(Y))
end

@testset "subspace_rr on an exact invariant subspace" begin
    n = 400
    A = lap1d(n)
    λall = lap1d_eigs(n)
    V = [sqrt(2 / (n + 1)) * sin(i * k * pi / (n + 1)) for i in 1:n, k in 191:210]
    Y = V * randn(20, 20)             # same span, mixed columns
    λ, X, res, Q = subspace_rr(A, Y)
    @test maximum(abs.(λ .- λall[191:210])) < 1e-12
    @test maximum(res) < 1e-12
    @test maximum(abs.(Q' * Q - I)) < 1e-12
end
