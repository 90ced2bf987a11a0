# ChainRulesCore reverse rule over libgodunov's discrete adjoint
# (gd_adjoint_run): the cotangent of the initial conservative state of N
# fixed-dt ns steps for the summed wall force.  Include after LibGodunov.jl,
# with ChainRulesCore in the environment:
#
#   include("LibGodunov.jl"); include("LibGodunovChainRules.jl")
#   using .LibGodunov, .LibGodunovChainRules, Zygote
#   r = GodunovRun(testMesh, params)            # a GdParams of the ns scheme
#   g = Zygote.gradient(U -> drag_sum(U, r, 200, dt), U0)[1]
module LibGodunovChainRules

using ChainRulesCore
using ..LibGodunov: GodunovRun, set_conservative!, run!, drag_gradient

export drag_sum

"""drag_sum(U0, r, N, dt; wall = -3, w = (1, 0)): the run's state set to U0
((n_cells, 4) column-major), advanced N steps at the fixed dt; returns
sum over the steps of w . F_wall, F_wall the pressure and viscous force on the
faces of boundary code `wall`.  The run ends at U_N."""
function drag_sum(U0::Matrix{Float64}, r::GodunovRun, N::Integer, dt::Real; wall = -3, w = (1.0, 0.0))
    set_conservative!(r, U0)
    s = run!(r, N; dt = dt, wall_code = wall)
    w[1] * s.wall_force_sum[1] + w[2] * s.wall_force_sum[2]
end

# The adjoint runs the forward sweep itself (keeping a checkpoint every
# `every` steps and the records of as many trailing segments as `budget`
# bytes hold), so the primal and its pullback come from one library call.
function ChainRulesCore.rrule(::typeof(drag_sum), U0::Matrix{Float64}, r::GodunovRun, N::Integer, dt::Real;
                              wall = -3, w = (1.0, 0.0), every = 20, budget = 0)
    set_conservative!(r, U0)
    g0, _, s = drag_gradient(r, N, dt; every = every, wall = wall, w = w, budget = budget)
    drag_sum_pullback(dJ) = (NoTangent(), unthunk(dJ) .* g0, NoTangent(), NoTangent(),
                             @not_implemented("dJ/ddt is not computed by the discrete adjoint"))
    return s.J, drag_sum_pullback
end

end  # module
