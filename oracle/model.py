"""Oracle-side local Hilbert space: numpy restatement of the reference's model setup.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module; it builds
the block structure the C oracle (``ctqmc_oracle.c``) samples with, following

* ``States.F90:58-184``   occupation basis, ``qns2substates`` blocking, offsets
* ``Operator.F90:115-200`` ``init_Psi``: c/c+ blocks and the ``Connect`` table
* ``Operator.F90:481-542, 1041-1109`` one-particle and U terms of H_loc
* ``Operator.F90:1111-1124`` ``force_hermitian``
* ``Operator.F90:637-670`` per-block ``DSYEV`` (here ``numpy.linalg.eigh``)
* ``Operator.F90:1130-1168`` ``transform_Psis``: eigenbasis operator blocks
* ``Trace.F90:559-568`` ``Egs`` = lowest eigenvalue (no outer truncation)
"""

from __future__ import annotations

import numpy as np


def _nt(state: int, n: int) -> int:
    return bin(state).count("1")


def _szt2(state: int, n: int) -> int:
    """2*Sz with the reference's convention: bits 0..N-1 count -1/2
    (``States.F90:224-234``)."""
    lo = bin(state & ((1 << n) - 1)).count("1")
    hi = bin(state >> n).count("1")
    return hi - lo


def _refine_blocks(states, h0, nstates: int, nb: int, thr: float):
    """analyze_hamiltonian (Operator.F90:303-442): within every block the
    connected components of H (|H_ij| > thr), numbered in order of discovery
    (breadth-first from each block's lowest unassigned state); then, while some
    operator maps two states of one block into two different blocks, those two
    blocks are merged (the larger label into the smaller, labels above it
    shifted down), rescanning from the first operator after every merge."""
    s2s = -np.ones(nstates, dtype=int)
    nsub = -1
    for sst, st in enumerate(states):
        n = len(st)
        H = h0[sst]
        nets = [[0]]
        nsub += 1
        s2s[st[0]] = nsub
        k = 0
        while k < len(nets):
            net = nets[k]
            q = 0
            while q < len(net):
                node = net[q]
                for edge in range(n):
                    v = H[node, edge]
                    if v != 0.0 and abs(v) > thr and edge not in net:
                        net.append(edge)
                        s2s[st[edge]] = nsub
                q += 1
            for node in range(n):
                if s2s[st[node]] == -1:
                    nets.append([node])
                    nsub += 1
                    s2s[st[node]] = nsub
                    break
            k += 1
    idx = np.arange(nstates)
    change = True
    while change:
        change = False
        for ib in range(nb):
            for sp in range(2):
                bit = ib + nb * sp
                for ca in (1, 2):
                    occ = (idx >> bit) & 1
                    cs = -np.ones(nstates, dtype=int)
                    ok = (occ == 0) if ca == 1 else (occ == 1)
                    cs[ok] = s2s[idx[ok] ^ (1 << bit)]
                    hit = None
                    for i1 in range(nstates):
                        if cs[i1] == -1:
                            continue
                        same = (s2s[i1 + 1:] == s2s[i1]) & (cs[i1 + 1:] != cs[i1]) & (cs[i1 + 1:] != -1)
                        j = np.flatnonzero(same)
                        if len(j):
                            hit = (i1, i1 + 1 + j[0])
                            break
                    if hit is not None:
                        a, b = cs[hit[0]], cs[hit[1]]
                        big, small = max(a, b), min(a, b)
                        s2s[s2s == big] = small
                        s2s[s2s > big] -= 1
                        change = True
                        break
                if change:
                    break
            if change:
                break
    return s2s, int(s2s.max()) + 1


class OracleModel:
    """Block-diagonal local Hamiltonian in its eigenbasis."""

    def __init__(self, umatrix: np.ndarray, muimp: np.ndarray, quantum_numbers: str = "Nt Szt",
                 epsblock: float = 1e-13):
        n = umatrix.shape[0]
        self.nbands = n
        nstates = 4 ** n
        self.nstates = nstates
        # qns2substates (States.F90:151-184): first-come order over state indices
        s2s = -np.ones(nstates, dtype=int)
        nsst = 0
        for st1 in range(nstates):
            if s2s[st1] != -1:
                continue
            s2s[st1] = nsst
            for st2 in range(st1, nstates):
                if s2s[st2] != -1:
                    continue
                if "Nt" in quantum_numbers and _nt(st1, n) != _nt(st2, n):
                    continue
                if "Szt" in quantum_numbers and _szt2(st1, n) != _szt2(st2, n):
                    continue
                s2s[st2] = nsst
            nsst += 1
        if "All" in quantum_numbers:
            # analyze_hamiltonian (CTQMC.F90:7971-7989, Operator.F90:303-442):
            # H on the blocks above, then its connectivity refines them
            self._set_blocks(s2s, nsst)
            h0 = self._hloc(umatrix, muimp, strict=False)
            h0 = [0.5 * (x + x.T) for x in h0]
            thr = epsblock * max(np.abs(muimp).max(), np.abs(umatrix).max())
            s2s, nsst = _refine_blocks(self.states, h0, s2s.size, n, thr)
        self._set_blocks(s2s, nsst)
        self.connect = self._connect
        # H_loc blocks
        h = self._hloc(umatrix, muimp, strict=False)
        self._finish(h)

    def _set_blocks(self, s2s, nsst):
        """Superstates from a state -> block map: states ascending within a block
        (init_SubStates, States.F90:98-146), Connect table and c/c+ blocks."""
        n = self.nbands
        self.nsst = nsst
        self.states = [np.where(s2s == s)[0] for s in range(nsst)]
        self.sizes = np.array([len(x) for x in self.states])
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)[:-1]])
        self.state_sst = s2s

        # c/c+ blocks in the occupation basis (Operator.F90:115-200)
        # connect[s, b, spin, ca] with ca 0 = creator, 1 = annihilator
        connect = -np.ones((nsst, n, 2, 2), dtype=int)
        psi = {}
        for s1 in range(nsst):
            for b in range(n):
                for sp in range(2):
                    bit = b + n * sp
                    for ca in range(2):
                        mat = {}
                        for i1, st in enumerate(self.states[s1]):
                            occ = (st >> bit) & 1
                            if (ca == 0 and occ) or (ca == 1 and not occ):
                                continue
                            t = st ^ (1 << bit)
                            sign = -1.0 if bin(st & ((1 << bit) - 1)).count("1") % 2 else 1.0
                            s2 = s2s[t]
                            mat.setdefault(s2, []).append((i1, t, sign))
                        if not mat:
                            continue
                        assert len(mat) == 1, "operator maps one block to several"
                        (s2, ent), = mat.items()
                        connect[s1, b, sp, ca] = s2
                        m = np.zeros((self.sizes[s2], self.sizes[s1]))
                        pos = {st: i for i, st in enumerate(self.states[s2])}
                        for i1, t, sign in ent:
                            m[pos[t], i1] = sign
                        psi[(s1, b, sp, ca)] = m
        self._connect = connect
        self.connect = connect
        self._psi = psi

    def _finish(self, h):
        self.hblocks = [0.5 * (x + x.T) for x in h]

        self.evals = []
        self.evecs = []
        for x in self.hblocks:
            w, v = np.linalg.eigh(x)
            self.evals.append(w)
            self.evecs.append(v)
        self.egs = min(w[0] for w in self.evals)

        # eigenbasis operator blocks
        self.psi_eb = {}
        for (s1, b, sp, ca), m in self._psi.items():
            s2 = self.connect[s1, b, sp, ca]
            self.psi_eb[(s1, b, sp, ca)] = self.evecs[s2].T @ m @ self.evecs[s1]
        self.psi_occ = self._psi

    # ------------------------------------------------------------------
    def _hloc(self, umatrix: np.ndarray, muimp: np.ndarray, strict: bool) -> list:
        """Blocks of 1/2 U_1234 c+_1 c+_2 c_4 c_3 + muimp_12 c+_1 c_2 in the
        occupation basis (``Operator.F90:481-542, 1041-1109``).  With
        ``strict`` a term that leaves a block raises (a derivative that breaks
        the quantum numbers)."""
        n, nsst, connect, psi = self.nbands, self.nsst, self.connect, self._psi
        h = [np.zeros((k, k)) for k in self.sizes]

        def leaves(s1, s):
            if strict and s1 >= 0 and s1 != s:
                raise ValueError("term maps superstate %d onto %d" % (s, s1))
            return s1 != s
        for b1 in range(n):
            for sp1 in range(2):
                for b2 in range(n):
                    for sp2 in range(2):
                        v = muimp[b1, sp1, b2, sp2]
                        if v == 0:
                            continue
                        for s in range(nsst):
                            s2 = connect[s, b2, sp2, 1]
                            if s2 < 0:
                                continue
                            s1 = connect[s2, b1, sp1, 0]
                            if leaves(s1, s):
                                continue
                            h[s] += v * psi[(s2, b1, sp1, 0)] @ psi[(s, b2, sp2, 1)]
        um = umatrix.reshape((n, 2) * 4)
        for idx in np.argwhere(um != 0):
            b1, sp1, b2, sp2, b3, sp3, b4, sp4 = idx
            v = 0.5 * um[tuple(idx)]
            for s in range(nsst):
                s4 = connect[s, b3, sp3, 1]
                if s4 < 0:
                    continue
                s3 = connect[s4, b4, sp4, 1]
                if s3 < 0:
                    continue
                s2 = connect[s3, b2, sp2, 0]
                if s2 < 0:
                    continue
                s1 = connect[s2, b1, sp1, 0]
                if leaves(s1, s):
                    continue
                h[s] += v * (psi[(s2, b1, sp1, 0)] @ psi[(s3, b2, sp2, 0)]
                             @ psi[(s4, b4, sp4, 1)] @ psi[(s, b3, sp3, 1)])
        return h

    def derivative_blocks(self, du: np.ndarray, dmuimp: np.ndarray) -> np.ndarray:
        """dH/dtheta in the eigenbasis, V_s^T (dH)_s V_s per superstate, for
        inputs linear in (U matrix, muimp): the blocks concatenated column-major
        at the density-matrix offsets (sum of n_s^2 entries)."""
        h = self._hloc(np.asarray(du), np.asarray(dmuimp), strict=True)
        out = []
        for s in range(self.nsst):
            x = 0.5 * (h[s] + h[s].T)
            out.append(np.asfortranarray(self.evecs[s].T @ x @ self.evecs[s]).ravel(order="F"))
        return np.concatenate(out)

    # ------------------------------------------------------------------
    def flat(self) -> dict:
        """Arrays handed to the C oracle (``oracle_create``)."""
        n = self.nbands
        nsst = self.nsst
        nmax = int(self.sizes.max())
        evals = np.zeros(int(self.sizes.sum()))
        for s in range(nsst):
            evals[self.offsets[s]:self.offsets[s] + self.sizes[s]] = self.evals[s]
        # operator blocks: for each (s, b, sp, ca) with a target, a dense
        # column-major (n_target x n_source) block; opoff = -1 if none
        opoff = -np.ones((nsst, n, 2, 2), dtype=np.int64)
        chunks = []
        pos = 0
        for s in range(nsst):
            for b in range(n):
                for sp in range(2):
                    for ca in range(2):
                        m = self.psi_eb.get((s, b, sp, ca))
                        if m is None:
                            continue
                        opoff[s, b, sp, ca] = pos
                        chunks.append(np.asfortranarray(m).ravel(order="F"))
                        pos += m.size
        ops = np.concatenate(chunks) if chunks else np.zeros(1)
        evec = np.zeros(int((self.sizes ** 2).sum()))
        evoff = np.zeros(nsst, dtype=np.int64)
        p = 0
        for s in range(nsst):
            evoff[s] = p
            evec[p:p + self.sizes[s] ** 2] = np.asfortranarray(self.evecs[s]).ravel(order="F")
            p += self.sizes[s] ** 2
        first_state = np.array([self.states[s][0] for s in range(nsst)], dtype=np.int32)
        return dict(nsst=nsst, nmax=nmax, sizes=self.sizes.astype(np.int32),
                    offsets=self.offsets.astype(np.int32), evals=evals, egs=self.egs,
                    connect=self.connect.astype(np.int32), opoff=opoff, ops=ops,
                    evec=evec, evoff=evoff, first_state=first_state,
                    state_sst=self.state_sst.astype(np.int32),
                    states=np.concatenate(self.states).astype(np.int32))
