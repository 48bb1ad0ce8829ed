"""Sparsity pattern + slot permutation of one element block.

``Pattern.build(conn, n_points)`` is the setup half of
``ElementAssembler.from_mesh`` (tensormesh/assemble/element_assembler.py:804-853):
the CSR of the sorted unique ``(conn[e,i], conn[e,j])`` pairs and, for each
entry, the ascending list of slots ``s = e*B*B + i*B + j`` that the
reference's ``ReduceProjector`` sums into it.  The integer work runs in
libtmhip's host pattern builder; results are moved to the mesh device once
and cached by the assembler.
"""

from __future__ import annotations

import ctypes
import hashlib
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _native as nat


def layout_hash(row: np.ndarray, col: np.ndarray) -> str:
    """SHA-256 of the row bytes followed by the col bytes, each at the dtype it
    comes in: the digest ``SparseMatrix.__init__`` computes from the tensors it
    is handed (tensormesh/sparse/matrix.py:81-86)."""
    return hashlib.sha256(np.ascontiguousarray(row).tobytes() + np.ascontiguousarray(col).tobytes()).hexdigest()


INT32_MAX = 2 ** 31 - 1


def edge_index_dtype(n_points: int, n_slots: int) -> np.dtype:
    """Index dtype of the edges ``ElementAssembler.from_mesh`` hands to
    ``SparseMatrix``: scipy's ``coo -> csr -> coo`` of the slot pairs
    (element_assembler.py:821-825, 863) keeps int32 unless the matrix side or
    the slot count (the COO entries before duplicates are summed) exceeds the
    int32 range, where scipy switches to int64."""
    return np.dtype(np.int32) if max(n_points, n_slots) <= INT32_MAX else np.dtype(np.int64)


def repeated_vertex(conn: torch.Tensor) -> bool:
    """True if some element lists a node twice (the row kernels stage one
    diagonal value per incidence, which assumes distinct vertices)."""
    c = conn.detach()
    B = c.shape[1]
    eq = (c.unsqueeze(2) == c.unsqueeze(1)).sum(dim=(1, 2))
    return bool((eq > B).any().item())


@dataclass
class Pattern:
    n_points: int
    n_elem: int
    n_basis: int
    nnz: int
    row_ptr: torch.Tensor      # int32 [N+1]
    col_idx: torch.Tensor      # int32 [nnz]
    seg_ptr: torch.Tensor      # int32 [nnz+1]
    seg_slot: torch.Tensor     # int32 [S]
    slot2nnz: torch.Tensor     # int32 [S]
    row: torch.Tensor          # int64 [nnz]
    col: torch.Tensor          # int64 [nnz]
    inc_ptr: Optional[torch.Tensor] = None   # int32 [N+1]
    inc: Optional[torch.Tensor] = None       # int32 [E*B]

    def __post_init__(self):
        self._plans = {}
        self._colpos = None
        self._hash = None

    @property
    def edge_dtype(self) -> torch.dtype:
        """torch dtype of the reference's scalar-output edges (``edge_index_dtype``)."""
        dt = edge_index_dtype(self.n_points, self.n_elem * self.n_basis * self.n_basis)
        return torch.int32 if dt == np.int32 else torch.int64

    @property
    def hash(self) -> str:
        """Layout digest of the scalar output, as the reference's
        ``SparseMatrix`` computes it from the scipy edges (int32 row bytes +
        int32 col bytes at every size the kernels accept; see
        ``edge_index_dtype``).  Computed on first use from ``row_ptr`` /
        ``col_idx`` and kept, so the 64 MB host round trip of the reference's
        per-construction hash is paid once per pattern, and only when a
        caller asks for the digest."""
        if self._hash is None:
            dt = edge_index_dtype(self.n_points, self.n_elem * self.n_basis * self.n_basis)
            rp = self.row_ptr.cpu().numpy()
            row = np.repeat(np.arange(self.n_points, dtype=dt), np.diff(rp))
            col = self.col_idx.cpu().numpy().astype(dt, copy=False)
            self._hash = layout_hash(row, col)
        return self._hash

    # blocks of the row-centric P1 kernels (tm_rows.hip): at most ROWS_CAP
    # incidences and ROWS_MAX rows per workgroup
    ROWS_CAP = 512
    ROWS_MAX = 64

    def colpos(self, conn: torch.Tensor) -> Optional[torch.Tensor]:
        """Packed per-incidence column positions for the fused P1 path, or
        None when the element has more than 4 nodes, a row has more than 255
        entries or an element repeats a vertex (the two-stage path then serves
        the call)."""
        if self._colpos is None:
            if self.n_basis > 4 or self.inc is None or repeated_vertex(conn):
                self._colpos = False
            else:
                c = np.ascontiguousarray(conn.detach().cpu().numpy().astype(np.int32))
                rp, ci = self.row_ptr.cpu().numpy(), self.col_idx.cpu().numpy()
                ip, ic = self.inc_ptr.cpu().numpy(), self.inc.cpu().numpy()
                out = np.empty(ic.shape[0], np.uint32)
                mask = np.empty(self.nnz, np.uint32)
                st = nat.lib().tm_colpos_build(c.ctypes.data, c.shape[0], c.shape[1], self.n_points,
                                               rp.ctypes.data, ci.ctypes.data, ip.ctypes.data, ic.ctypes.data,
                                               out.ctypes.data, mask.ctypes.data)
                if st == 0:
                    self._colpos = torch.from_numpy(out.view(np.int32)).to(self.seg_ptr.device)
                    self._entry_mask = torch.from_numpy(mask.view(np.int32)).to(self.seg_ptr.device)
                else:
                    self._colpos = False
        return self._colpos if self._colpos is not False else None

    def inc_inv(self) -> torch.Tensor:
        """Position of incidence e*B + i in the node-sorted incidence list (int32)."""
        if getattr(self, "_inc_inv", None) is None:
            n = self.inc.numel()
            inv = torch.empty(n, dtype=torch.int32, device=self.inc.device)
            inv[self.inc.long()] = torch.arange(n, dtype=torch.int32, device=self.inc.device)
            self._inc_inv = inv
        return self._inc_inv

    def rows_view(self, conn: torch.Tensor, order: bool = True) -> Optional[nat.RowsView]:
        """ctypes view of the fused P1 path, or None when it cannot serve this
        pattern.  ``order``: with the contribution order (``rows_order``) on
        a device pattern, so every entry is a contiguous sum in the kernel."""
        cp = self.colpos(conn)
        if cp is None:
            return None
        bp = self.plan(self.ROWS_CAP, node=True, max_rows=self.ROWS_MAX)
        info = self.block_info(self.ROWS_CAP)
        rv = nat.RowsView(self.nnz, self.row_ptr.data_ptr(), self.inc_ptr.data_ptr(), self.inc.data_ptr(),
                          cp.data_ptr(), self._entry_mask.data_ptr(), bp.numel() - 1, bp.data_ptr(),
                          info.data_ptr(), None, None)
        if order and self.row_ptr.device.type == "cuda":
            dest, ent_start = self.rows_order(rv)
            rv.dest, rv.ent_start = dest.data_ptr(), ent_start.data_ptr()
        return rv

    def rows_order(self, rv: nat.RowsView):
        """(dest [E*B, 4] int16, ent_start [nnz] int16) of tm_rows_dest_build for
        the fused path's plan, built once on the pattern's device."""
        if getattr(self, "_rows_order", None) is None:
            dev = self.row_ptr.device
            dest = torch.zeros(self.inc.numel() * 4, dtype=torch.int16, device=dev)
            ent_start = torch.zeros(max(self.nnz, 1), dtype=torch.int16, device=dev)
            nat.check(nat.lib().tm_rows_dest_build(self.n_basis, ctypes.byref(rv), dest.data_ptr(),
                                                   ent_start.data_ptr(), nat.stream_of(dev)), "tm_rows_dest_build")
            self._rows_order = (dest, ent_start)
        return self._rows_order

    def block_info(self, cap: int) -> torch.Tensor:
        """[n_blocks + 1, 4] int32: first row, first incidence and first CSR
        entry of every block of the node plan (one 16-byte load per bound)."""
        key = ("info", cap)
        if key not in self._plans:
            bp = self.plan(cap, node=True, max_rows=self.ROWS_MAX).long()
            info = torch.zeros(bp.numel(), 4, dtype=torch.int32, device=bp.device)
            info[:, 0] = bp.to(torch.int32)
            info[:, 1] = self.inc_ptr[bp]
            info[:, 2] = self.row_ptr[bp]
            self._plans[key] = info.contiguous()
        return self._plans[key]

    def plan(self, cap: int, node: bool = False, max_rows: Optional[int] = None) -> torch.Tensor:
        """Reduction plan (blk_ptr, int32, on the pattern's device) for entries
        (node=False) or for the node incidence lists (node=True); max_rows
        caps the segments (rows) per block."""
        key = (cap, node, max_rows)
        if key not in self._plans:
            seg = (self.inc_ptr if node else self.seg_ptr).cpu().numpy()
            self._plans[key] = torch.from_numpy(nat.reduce_plan(seg, cap, max_rows)).to(self.seg_ptr.device)
        return self._plans[key]

    @staticmethod
    def build(conn: torch.Tensor, n_points: int, device=None, incidence: bool = True) -> "Pattern":
        L = nat.lib()
        c = np.ascontiguousarray(conn.detach().cpu().numpy().astype(np.int32, copy=False))
        E, B = c.shape
        nnz = ctypes.c_int64(0)
        nat.check(L.tm_pattern_nnz(c.ctypes.data, E, B, n_points, ctypes.byref(nnz)), "tm_pattern_nnz")
        nnz = int(nnz.value)
        S = E * B * B
        row_ptr = np.empty(n_points + 1, np.int32)
        col_idx = np.empty(nnz, np.int32)
        seg_ptr = np.empty(nnz + 1, np.int32)
        seg_slot = np.empty(S, np.int32)
        slot2nnz = np.empty(S, np.int32)
        nat.check(L.tm_pattern_build(c.ctypes.data, E, B, n_points, row_ptr.ctypes.data, col_idx.ctypes.data,
                                     seg_ptr.ctypes.data, seg_slot.ctypes.data, slot2nnz.ctypes.data),
                  "tm_pattern_build")
        row = np.repeat(np.arange(n_points, dtype=np.int64), np.diff(row_ptr))
        col = col_idx.astype(np.int64)
        inc_ptr = inc = None
        if incidence:
            inc_ptr_np = np.empty(n_points + 1, np.int32)
            inc_np = np.empty(E * B, np.int32)
            nat.check(L.tm_incidence_build(c.ctypes.data, E, B, n_points, inc_ptr_np.ctypes.data,
                                           inc_np.ctypes.data), "tm_incidence_build")
            inc_ptr, inc = torch.from_numpy(inc_ptr_np), torch.from_numpy(inc_np)
        dev = conn.device if device is None else torch.device(device)
        t = lambda a: torch.from_numpy(a).to(dev)
        return Pattern(n_points, E, B, nnz, t(row_ptr), t(col_idx), t(seg_ptr), t(seg_slot), t(slot2nnz),
                       t(row), t(col),
                       None if inc_ptr is None else inc_ptr.to(dev), None if inc is None else inc.to(dev))

    @staticmethod
    def build_device(conn: torch.Tensor, n_points: int) -> Optional["Pattern"]:
        """The same pattern built on the GPU (tm_pattern_device_count / _fill).
        Returns None when a row has too many neighbour keys for the device
        builder (the host builder then serves)."""
        L = nat.lib()
        dev = conn.device
        c = conn.to(torch.int32).contiguous()
        E, B = c.shape
        if E == 0:
            return None   # no elements: the host builder gives the empty pattern
        st = nat.stream_of(dev)
        i32 = lambda n: torch.empty(n, dtype=torch.int32, device=dev)
        inc_ptr, inc, row_ptr, work = i32(n_points + 1), i32(E * B), i32(n_points + 1), i32(n_points + 1)
        nat.check(L.tm_pattern_device_count(c.data_ptr(), E, B, n_points, inc_ptr.data_ptr(), inc.data_ptr(),
                                            row_ptr.data_ptr(), work.data_ptr(), st), "tm_pattern_device_count")
        nnz = int(row_ptr[n_points].item())
        if int(work[n_points].item()) != 0:
            return None
        S = E * B * B
        col_idx, seg_ptr, seg_slot, slot2nnz = i32(nnz), i32(nnz + 1), i32(S), i32(S)
        colpos = i32(E * B) if B <= 4 else None
        mask = i32(nnz) if B <= 4 else None
        p = lambda t: None if t is None else t.data_ptr()
        nat.check(L.tm_pattern_device_fill(c.data_ptr(), E, B, n_points, inc_ptr.data_ptr(), inc.data_ptr(),
                                           row_ptr.data_ptr(), col_idx.data_ptr(), seg_ptr.data_ptr(),
                                           seg_slot.data_ptr(), slot2nnz.data_ptr(), p(colpos), p(mask), st),
                  "tm_pattern_device_fill")
        row = torch.repeat_interleave(torch.arange(n_points, device=dev), (row_ptr[1:] - row_ptr[:-1]).long())
        col = col_idx.long()
        pat = Pattern(n_points, E, B, nnz, row_ptr, col_idx, seg_ptr, seg_slot, slot2nnz, row, col, inc_ptr, inc)
        if colpos is not None:
            too_long = bool(((row_ptr[1:] - row_ptr[:-1]) > 255).any().item())
            if not too_long and not repeated_vertex(conn):
                pat._colpos, pat._entry_mask = colpos, mask
            else:
                pat._colpos = False
        return pat

    def to(self, device) -> "Pattern":
        dev = torch.device(device)
        mv = lambda x: None if x is None else x.to(dev)
        p = Pattern(self.n_points, self.n_elem, self.n_basis, self.nnz, mv(self.row_ptr), mv(self.col_idx),
                    mv(self.seg_ptr), mv(self.seg_slot), mv(self.slot2nnz), mv(self.row), mv(self.col),
                    mv(self.inc_ptr), mv(self.inc))
        p._hash = self._hash
        return p

    def sym_slot(self) -> torch.Tensor:
        """Segment list for symmetric element blocks (TM_FLAG_UPPER): every slot
        renamed to its packed upper block (the element's pairs i <= j stored
        contiguously), with bit 31 set where the pair has i > j (read
        transposed); built once on the pattern's device by tm_symmetric_segments."""
        if getattr(self, "_sym_slot", None) is None:
            nat.require_device(self.seg_slot, "pattern")
            sym = torch.empty_like(self.seg_slot)
            nat.check(nat.lib().tm_symmetric_segments(self.n_basis, self.seg_slot.numel(), self.seg_slot.data_ptr(),
                                                      sym.data_ptr(), nat.stream_of(self.seg_slot.device)),
                      "tm_symmetric_segments")
            self._sym_slot = sym
        return self._sym_slot

    def sym_pairs(self):
        """Paired segment list for symmetric element blocks: the upper entries
        (row <= column) in CSR order, each with its contributions from
        ``sym_slot`` (ascending element order, as in ``seg_slot``) and the
        position of its mirror entry (column, row).  Returns device int32
        tensors (entry, entry_t, seg_ptr, seg_slot); built once."""
        if getattr(self, "_sym_pairs", None) is None:
            dev = self.seg_slot.device
            rp = self.row_ptr.cpu().numpy().astype(np.int64)
            ci = self.col_idx.cpu().numpy().astype(np.int64)
            sp = self.seg_ptr.cpu().numpy().astype(np.int64)
            sym = self.sym_slot().cpu().numpy()
            N = self.n_points
            rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(rp))
            up = np.flatnonzero(rows <= ci)
            keys = rows * N + ci                      # ascending: CSR order
            mirror = ci[up] * N + rows[up]
            ent_t = np.searchsorted(keys, mirror)
            if not np.array_equal(keys[np.minimum(ent_t, keys.size - 1)], mirror):
                raise ValueError("element pattern is not structurally symmetric")
            lens = sp[up + 1] - sp[up]
            seg_ptr = np.zeros(up.size + 1, np.int64)
            np.cumsum(lens, out=seg_ptr[1:])
            idx = np.repeat(sp[up] - seg_ptr[:-1], lens) + np.arange(seg_ptr[-1], dtype=np.int64)
            t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)
            self._sym_pairs = (t(up), t(ent_t), t(seg_ptr), t(sym[idx]))
            self._sym_plans = {}
            self._sym_slot = None   # only the paired list is read from here on
        return self._sym_pairs

    #: symmetric element blocks are written in contribution order (slot_dest)
    #: so that the paired reduction streams contiguous slots; False keeps the
    #: element-ordered packed blocks and gathers them through the segment list
    contiguous_slots: bool = True

    def slot_dest(self) -> Optional[torch.Tensor]:
        """Contribution-ordered slot of every packed upper block (int32 [E*P],
        bit 31 = stored transposed), the inverse of the paired segment list;
        None when an element repeats a vertex.  Every packed block is listed
        at least once among the upper entries' contributions, so a list of
        exactly E*P contributions lists each once (a repeated vertex lists
        some twice, on the diagonal).  Built once on the pattern's device."""
        if getattr(self, "_slot_dest", None) is None:
            _, _, _, seg_slot = self.sym_pairs()
            B = self.n_basis
            n = self.n_elem * (B * (B + 1) // 2)
            if seg_slot.numel() != n:
                self._slot_dest = False
            else:
                dest = torch.empty(n, dtype=torch.int32, device=seg_slot.device)
                nat.check(nat.lib().tm_slot_dest_build(n, seg_slot.data_ptr(), dest.data_ptr(),
                                                       nat.stream_of(seg_slot.device)), "tm_slot_dest_build")
                self._slot_dest = dest
        return self._slot_dest if self._slot_dest is not False else None

    def view(self, dtype: torch.dtype = torch.float64, hh: int = 1, symmetric: bool = False) -> nat.PatternView:
        """ctypes view for the reductions; ``symmetric`` selects the paired
        upper-entry list of ``sym_pairs`` (element kernels launched with
        TM_FLAG_UPPER): each upper entry is summed once and mirrored, and
        (``contiguous_slots``) its contributions are consecutive slots that
        the element kernels write through ``slot_dest``."""
        if not symmetric:
            bp = self.plan(nat.reduce_cap(dtype, hh))
            return nat.PatternView(self.nnz, self.row_ptr.data_ptr(), self.col_idx.data_ptr(),
                                   self.seg_ptr.data_ptr(), self.seg_slot.data_ptr(), bp.numel() - 1, bp.data_ptr())
        entry, entry_t, seg_ptr, seg_slot = self.sym_pairs()
        cap = nat.reduce_cap(dtype, hh)
        if cap not in self._sym_plans:
            self._sym_plans[cap] = torch.from_numpy(nat.reduce_plan(seg_ptr.cpu().numpy(), cap)).to(seg_ptr.device)
        bp = self._sym_plans[cap]
        sd = self.slot_dest() if self.contiguous_slots else None
        return nat.PatternView(entry.numel(), self.row_ptr.data_ptr(), self.col_idx.data_ptr(), seg_ptr.data_ptr(),
                               seg_slot.data_ptr(), bp.numel() - 1, bp.data_ptr(), entry.data_ptr(),
                               entry_t.data_ptr(), None if sd is None else sd.data_ptr())
