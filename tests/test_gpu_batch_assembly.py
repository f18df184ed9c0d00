"""Batched assembly of many trilinear meshes (mife_assemble_batch) against each
mesh's own assembly call, bit for bit, in FP64 and FP32, with and without a
source vector, for scalar (diffusion) and elasticity meshes of different sizes in
one batch."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


class _Mesh:
    """A generated trilinear box with the generic structures: CSR, position map,
    structured colouring and the per-mesh assembler."""

    def __init__(self, kind, ne):
        from iron_amd import fem
        from iron_amd.basis import basis
        from iron_amd.mesh import DofMap, RegularMesh, structured_colours
        self.kind = kind
        b = basis(3, 1, 2)
        m = RegularMesh(ne, [b], (1.0, 1.0, 1.0) if kind == "diffusion" else (3.0, 1.0, 1.0))
        self.mesh = m
        dm = DofMap(m, [0] * (1 if kind == "diffusion" else 3))
        self.csr = fem.CSRStructure(dm.element_dofs(), dm.num_dofs, "cuda", colours=structured_colours(m))
        A = fem.ScalarAssembler if kind == "diffusion" else fem.ElasticAssembler
        self.asm = A(b, m.coords, m.elem_nodes[0], self.csr)

    def coefficients(self, seed, dev):
        from iron_amd import serving
        z = serving.lognormal_field(self.mesh.num_elements, seed, dev)
        if self.kind == "diffusion":
            return z
        nu = torch.full_like(z, 0.3)
        return torch.stack([10 * z, 10 * z, 10 * z, nu, nu, nu], dim=1).contiguous()


def test_batch_equals_per_mesh_assembly():
    from iron_amd import _lib, fem
    from iron_amd.basis import basis
    print("case batched assembly of 6 meshes", flush=True)
    shapes = [("diffusion", (4, 3, 5)), ("elasticity", (6, 2, 3)), ("diffusion", (7, 7, 6)),
              ("elasticity", (3, 3, 3)), ("diffusion", (1, 1, 1)), ("elasticity", (1, 2, 1))]
    sts = [_Mesh(k, ne) for k, ne in shapes]
    q1 = fem.DeviceBasis(basis(3, 1, 2), "cuda")
    L = _lib.lib()
    for prec in (fem.F64, fem.F32):
        coefs = [st.coefficients(100 + i, "cuda") for i, st in enumerate(sts)]
        g = torch.Generator(device="cuda").manual_seed(7)
        srcs = [torch.rand(st.mesh.num_elements, dtype=torch.float64, device="cuda", generator=g)
                if st.kind == "diffusion" and i % 2 == 0 else None for i, st in enumerate(sts)]
        ref, ref_s, out, out_s = [], [], [], []
        for st, cf, sr in zip(sts, coefs, srcs):
            v = st.csr.zeros(prec)
            sv = torch.zeros(st.csr.n, dtype=torch.float64, device="cuda") if sr is not None else None
            if st.kind == "diffusion":
                st.asm.assemble(v, kappa=cf, source=sr, src_vec=sv, prec=prec)
            else:
                st.asm.assemble(v, cf, prec=prec)
            ref.append(v)
            ref_s.append(sv)
            out.append(st.csr.zeros(prec))
            out_s.append(torch.zeros(st.csr.n, dtype=torch.float64, device="cuda") if sr is not None else None)
        jobs = (_lib.AssemblyJob * len(sts))()
        maxc = 0
        for k, st in enumerate(sts):
            cs, asm = st.csr, st.asm
            kind = _lib.MIFE_JOB_SCALAR if st.kind == "diffusion" else _lib.MIFE_JOB_ELASTIC
            jobs[k] = _lib.AssemblyJob(kind, asm.n_elem, _lib.ptr(asm.coords), _lib.ptr(asm.elem_nodes),
                                       _lib.ptr(cs.elem_dofs), _lib.ptr(coefs[k]), _lib.ptr(srcs[k]), _lib.ptr(cs.pos),
                                       _lib.ptr(cs.order), cs.colour_ptr.ctypes.data, cs.n_colours, 0,
                                       _lib.ptr(out[k]), _lib.ptr(out_s[k]))
            maxc = max(maxc, cs.n_colours)
        nb = int(L.mife_assemble_batch_work_bytes(len(sts), maxc))
        host = np.empty(nb, dtype=np.uint8)
        dev = torch.empty(nb, dtype=torch.uint8, device="cuda")
        d = C.byref(q1.desc)
        _lib.call("mife_assemble_batch", d, d, len(sts), jobs, maxc, prec, host.ctypes.data, _lib.ptr(dev),
                  _lib.stream_ptr())
        torch.cuda.synchronize()
        for k in range(len(sts)):
            assert torch.count_nonzero(ref[k]) > 0
            assert torch.equal(ref[k], out[k]), (prec, shapes[k])
            if ref_s[k] is not None:
                assert torch.equal(ref_s[k], out_s[k]), (prec, shapes[k])
