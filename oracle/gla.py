"""GLA restatement (src/gla.rs, src/util.rs:142-172, src/ising.rs:327-355) in plain Python over the
oracle's instance arrays.

Test infrastructure: the product package never imports this module.
"""
import numpy as np


def read_u32_lines(text):
    """read_u32_lines (src/util.rs:142-172): per line many1(delimited(space0, u32, multispace0));
    lines whose first token is not an integer are skipped; the rest of a line after the last
    parsed integer is ignored."""
    out = []
    for line in text.split("\n"):
        if line.endswith("\r"):
            line = line[:-1]
        vals, p = [], 0
        while True:
            q = p
            while q < len(line) and line[q] in " \t":
                q += 1
            r = q
            while r < len(line) and line[r].isdigit() and line[r] in "0123456789":
                r += 1
            if r == q:
                break
            vals.append(int(line[q:r]))
            while r < len(line) and line[r] in " \t\r\n":
                r += 1
            p = r
        if vals:
            out.append(vals)
    return out


def read_graph_partition(text, n):
    """GlaParams::read_graph_partition (src/gla.rs:35-59) plus the sort of instance_partition
    (src/gla.rs:88-91).  Raises ValueError where the reference returns an error."""
    seen = np.zeros(n, bool)
    parts = read_u32_lines(text)
    for ip, p in enumerate(parts):
        for i in p:
            if i >= n:
                raise ValueError("Partitition must be specified for instance size %d. Found index %d" % (n, i))
            if seen[i]:
                raise ValueError("Index %d specified more than once in partition %d" % (i, ip))
            seen[i] = True
    return [sorted(p) for p in parts]


def induced_subgraph(row_ptr, col, J, bias, nodes):
    """induced_subgraph_instance_sorted (src/ising.rs:327-355) on coupling rows in CSR form:
    returns (row_ptr, col, J, bias) of the sub-instance; its ICM graph has the same pattern."""
    pos = {int(v): k for k, v in enumerate(nodes)}
    rp, cc, jj = [0], [], []
    for ni in nodes:
        for e in range(row_ptr[ni], row_ptr[ni + 1]):
            j0 = int(col[e])
            if j0 in pos:
                cc.append(pos[j0])
                jj.append(J[e])
        rp.append(len(cc))
    return (np.array(rp, np.uint32), np.array(cc, np.uint32), np.array(jj, np.float32),
            np.array([bias[ni] for ni in nodes], np.float32))


def partition_graph(row_ptr, col, J, parts):
    """instance_partition (src/gla.rs:73-120): edges (i, j, K) of the energy graph (coupling pattern,
    i < j, row order) between different partitions, grouped per directed (source partition, target
    partition) pair in order of first appearance.  Returns [((ps, pt), [(i, j, K), ...]), ...]."""
    part_of = {}
    for p, sites in enumerate(parts):
        for i in sites:
            part_of[int(i)] = p
    groups, order = {}, []
    n = len(row_ptr) - 1
    for i in range(n):
        for e in range(row_ptr[i], row_ptr[i + 1]):
            j = int(col[e])
            if j <= i:
                continue
            ps, pt = part_of[i], part_of[j]
            if ps == pt:
                continue
            if (ps, pt) not in groups:
                groups[(ps, pt)] = []
                order.append((ps, pt))
            groups[(ps, pt)].append((i, j, np.float32(J[e])))
    return [(k, groups[k]) for k in order]


def boundary_energies(row_ptr, col, J, parts, joined):
    """The boundary energies of run_gla (src/gla.rs:175-190) for the joined state: per partition
    edge, bnd += K * (s_i s_j) in f32 edge by edge; returns ([(ps, pt, bnd)], total in f32)."""
    out = []
    total = np.float32(0.0)
    for (ps, pt), es in partition_graph(row_ptr, col, J, parts):
        bnd = np.float32(0.0)
        for i, j, K in es:
            bnd = np.float32(bnd + np.float32(K * np.float32(int(joined[i]) * int(joined[j]))))
        total = np.float32(total + bnd)
        out.append((ps, pt, bnd))
    return out, total
