// Host-side numerics shared by the library's set-up code: Gauss-Legendre
// rules and the adaptive quadrature used for orientation-distribution
// coefficients and shape-distribution limits.
#pragma once

namespace ctm {

// Gauss-Legendre nodes (ascending) and weights on [-1, 1]
// (SpecialFunctions.hpp:647-715: Lether-Wenston start, Newton on the
// three-term recurrence, tolerance 1e-16 relaxed tenfold per iteration
// after the 100th).
void gauss_legendre_host(unsigned order, double *x, double *w);

// Nodes and weights on [a, b] (SpecialFunctions.hpp:738-751).
void gauss_legendre_host_ab(unsigned order, double a, double b, double *x,
                            double *w);

}  // namespace ctm
