"""CPU oracle for the tensorized element-assembly path.

TEST INFRASTRUCTURE ONLY.  Nothing under ``tensormesh_amd/`` imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker (or the
timed CPU baseline), never as the thing that is measured or shipped.

It restates the reference's algorithm (kilickursat/TensorMesh, pure
Python/PyTorch) op by op in float64 on the CPU: numpy/scipy for the integer
pattern work, torch-CPU for the floating-point chain (einsum, batched
inverse/det, ``index_add_``, autograd), following the reference files cited
in each function's docstring.

Pinning (what each function is checked against, besides the HIP path):

==========================  =================================================
function                    pins (tests/test_oracle*.py)
==========================  =================================================
pattern                     brute-force python set on small meshes; the
                            exact nnz counts 2,921 (20x20 tri) and
                            4,018,753 (64^3 Kuhn tets); slot->nnz round trip
quadrature / basis          weights sum to the reference volume; exactness
                            on monomials; Kronecker property; P2 triangle
                            values of tests/element/test_transformation.py
shape_grad_jxw              |det J| = 2*area (shoelace) / 6*volume (triple
                            product); gradients of linear functions exact
bilinear (laplace/diff.)    patch test K @ (a.x+b) = 0 at interior nodes;
                            row sums 0; symmetry; PSD; closed-form P1 element
                            matrix vol * grad(l_i).grad(l_j)
bilinear (mass)             1^T M 1 = |Omega|; closed-form P1 mass
                            |T|/((D+1)(D+2)) (1 + delta_ij)
bilinear (elasticity)       rigid-body modes in the kernel; Voigt B/C known
                            answers of tests/functional/test_elasticity.py
neo-hookean                 energy 0 at u=0; tangent(u=0) == linear
                            elasticity; residual == FD of energy; tangent ==
                            FD of residual
reduce_project              brute-force loop (NaiveProjector semantics of
                            tests/assemble/test_projector.py)
block_coo                   tests/sparse/test_block_coo.py naive example
node_assemble (load)        sum b = integral of f_h; closed-form P1 load
facet loads                 Gauss-Legendre exactness on monomials; facet
                            selection against plain loops; unit square:
                            sum b = 4 + 4 sqrt(2) h (f = 1), 2 + 2 sqrt(2) h
                            (f = x), the two corner diagonals included as the
                            reference selects them; one edge: h per inner
                            node, h/2 at its ends; P2 Simpson weights; P2
                            face: vertex functions integrate to 0; cube: sum
                            of selected triangle areas by cross products;
                            face x = 0: area 1, integral of y = 1/2
condense / recover          tests/operator/test_condenser.py 4x4 system
adjoint gradients           central finite differences; dense solve autograd
j2 (psi / derivatives /     yield stress -> inf: energy and Hessian equal
update)                     linear elasticity (2-D plane strain and 3-D);
                            simple shear closed form below and past yield;
                            residual / tangent == central FD past yield; the
                            radial return lands on the updated yield surface,
                            keeps eps_p trace-free, alpha grows by
                            sqrt(2/3)|d eps_p|, elastic points unchanged
contact (penalty)           per-facet adaptive quadrature (scipy quad /
                            dblquad) on straight facets; zero when clear of
                            the plane; gradient / Hessian == central FD
quadrilaterals, hexahedra   tensor rule exact on x^a y^b z^c, a,b,c <= 2n-1;
                            Kronecker / partition of unity; Q1 unit-cell
                            stiffness and mass as Kronecker products of the
                            1-D matrices; patch test on jittered grids;
                            rigid-body modes; boundary integrals 4 / 6 and 6 /
                            18 (f = 1 / f = x + 2y (+ 3z))
cubic simplices             Kronecker / partition of unity; patch test and
                            rigid modes on jittered P3 meshes; degree-6 rule
                            integrates the P3 mass to the cell volume; the
                            reference's rules reach the degrees recorded in
                            tests (tetrahedron "order 5" / "order 6" reach
                            degrees 4 / 5)
==========================  =================================================

Every function above has at least one such pin; none is "parity unpinned".
"""
