"""The reference's on-disk formats (the Python twin of host/gales_mc_io.hpp).

Text mesh (src/fem/mesh/mesh.hpp:141-456):

    MESH! 2D                                  (or 3D)
    nodes <N>
    elements <E>
    sides <S>
    X <min> <max>                             (one line per dimension)
    Y <min> <max>
    Node <gid> <x> <y> [<z>] <flag>                   N lines
    Element <gid> <pid> <nnodes> <n_0> ... <flag>     E lines (flag > 0: boundary)
    Side <gid> <pid> <nsn> <n_0> ... <flag>           S lines

Tokens are separated by spaces or commas (algorithm.hpp:219-225).  A side is
matched to the local side of a boundary element whose node set equals its own
(mesh.hpp:420-446, local numbering of element.hpp:145-149 / 455-459); sides
with flag 0 carry no boundary integral.  read_mesh(path, pid) keeps one
partition's elements and renumbers their nodes compactly (node_gid holds the
file gid of each local node).

Restart dumps (src/fem/base/io.hpp:84-181): the global state vector (dof gid =
node gid * ndof + k) as raw native float64 in "results/<dir><t>", t printed as
an ostream prints a double.

Magma solubility tables (mixtures.hpp:416-490), in "input/lookuptable/":
pressure.txt (a count, then that many pressures), an optional temperature.txt
(same layout), and per magma name wf_l_h2o_onliq_<name>.txt and
wf_l_co2_onliq_<name>.txt: n_p values, or n_T rows of n_p values with a
temperature table.  Values are whitespace separated.
"""
import re

import numpy as np

SIDE_TABLE = {2: [(1, 2), (2, 0), (0, 1)], 3: [(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)]}
_SPLIT = re.compile(r"[ ,\t\r\n]+")


def _lines(path):
    with open(path) as f:
        for line in f:
            t = [x for x in _SPLIT.split(line.strip()) if x]
            if t:
                yield t


def _expect(it, what):
    try:
        return next(it)
    except StopIteration:
        raise ValueError(f"mesh file ends before {what}") from None


def read_mesh(path, pid=None):
    """Mesh dict (dim, coords, conn, node_flag, side_elem, side_local,
    side_flag, node_gid, elem_gid, elem_pid) of a GALES text mesh."""
    it = _lines(path)
    t = _expect(it, "the header")
    if len(t) < 2 or t[0] != "MESH!" or t[1] not in ("2D", "3D"):
        raise ValueError("not a GALES mesh file (first line 'MESH! 2D' or 'MESH! 3D')")
    dim = 2 if t[1] == "2D" else 3
    nen = dim + 1
    counts = []
    for name in ("nodes", "elements", "sides"):
        t = _expect(it, name)
        if len(t) < 2 or t[0] != name:
            raise ValueError(f"expected '{name} <n>'")
        counts.append(int(t[1]))
    N, E, S = counts
    for _ in range(dim):
        _expect(it, "the extent lines")
    X = np.zeros((N, dim))
    flag = np.zeros(N, np.int32)
    seen = np.zeros(N, bool)
    for _ in range(N):
        t = _expect(it, "a Node line")
        if t[0] != "Node" or len(t) < dim + 3:
            raise ValueError("expected 'Node gid x y [z] flag'")
        g = int(t[1])
        if not 0 <= g < N or seen[g]:
            raise ValueError("Node gid out of range or repeated")
        seen[g] = True
        X[g] = [float(x) for x in t[2:2 + dim]]
        flag[g] = int(t[2 + dim])
    conn, egid, epid, eflag = [], [], [], []
    for _ in range(E):
        t = _expect(it, "an Element line")
        if t[0] != "Element" or len(t) < 4:
            raise ValueError("expected 'Element gid pid n nodes... flag'")
        nn = int(t[3])
        if nn != nen or len(t) < 5 + nn:
            raise ValueError(f"Element with {nn} nodes (linear triangles / tetrahedra only)")
        p = int(t[2])
        if pid is not None and p != pid:
            continue
        nodes = [int(x) for x in t[4:4 + nn]]
        if min(nodes) < 0 or max(nodes) >= N:
            raise ValueError("Element node out of range")
        egid.append(int(t[1]))
        epid.append(p)
        conn.append(nodes)
        eflag.append(int(t[4 + nn]))
    conn = np.array(conn, np.int64).reshape(-1, nen)
    used = np.arange(N) if pid is None else np.unique(conn)
    local = np.full(N, -1, np.int64)
    local[used] = np.arange(used.size)
    where = {}
    for e, en in enumerate(conn):
        if eflag[e] <= 0:
            continue
        for s, side in enumerate(SIDE_TABLE[dim]):
            key = tuple(sorted(int(en[a]) for a in side))
            if all(flag[g] > 0 for g in key):
                where.setdefault(key, (e, s))
    se, sl, sf = [], [], []
    for _ in range(S):
        t = _expect(it, "a Side line")
        if t[0] != "Side" or len(t) < 4:
            raise ValueError("expected 'Side gid pid n nodes... flag'")
        nsn = int(t[3])
        if nsn != dim or len(t) < 5 + nsn:
            raise ValueError("Side with the wrong node count")
        if pid is not None and int(t[2]) != pid:
            continue
        key = tuple(sorted(int(x) for x in t[4:4 + nsn]))
        f = int(t[4 + nsn])
        if key in where and f > 0:
            e, s = where[key]
            se.append(e)
            sl.append(s)
            sf.append(f)
    return dict(dim=dim, coords=np.ascontiguousarray(X[used]),
                conn=np.ascontiguousarray(local[conn], np.int32),
                node_flag=np.ascontiguousarray(flag[used]), node_gid=used.astype(np.int32),
                elem_gid=np.array(egid, np.int32), elem_pid=np.array(epid, np.int32),
                side_elem=np.array(se, np.int32), side_local=np.array(sl, np.int32),
                side_flag=np.array(sf, np.int32))


def write_mesh(path, mesh, elem_pid=None):
    """Writes a mesh dict in the text format (pid 0 unless elem_pid is given;
    an element is flagged when it carries a side)."""
    dim = mesh["dim"]
    X = np.asarray(mesh["coords"]).reshape(-1, dim)
    conn = np.asarray(mesh["conn"])
    nen = dim + 1
    n_s = len(mesh["side_elem"])
    pid = np.zeros(conn.shape[0], int) if elem_pid is None else np.asarray(elem_pid)
    on_bd = np.zeros(conn.shape[0], int)
    on_bd[np.asarray(mesh["side_elem"])] = 1
    with open(path, "w") as f:
        f.write(f"MESH! {dim}D\nnodes {X.shape[0]}\nelements {conn.shape[0]}\nsides {n_s}\n")
        for d in range(dim):
            f.write(f"{'XYZ'[d]} {X[:, d].min()!r} {X[:, d].max()!r}\n")
        for i in range(X.shape[0]):
            f.write("Node %d %s %d\n" % (i, " ".join(repr(float(x)) for x in X[i]), mesh["node_flag"][i]))
        for e in range(conn.shape[0]):
            f.write("Element %d %d %d %s %d\n" % (e, pid[e], nen, " ".join(str(int(a)) for a in conn[e]),
                                                   on_bd[e]))
        for s in range(n_s):
            e = int(mesh["side_elem"][s])
            nodes = [int(conn[e, a]) for a in SIDE_TABLE[dim][int(mesh["side_local"][s])]]
            f.write("Side %d %d %d %s %d\n" % (s, pid[e], dim, " ".join(map(str, nodes)), mesh["side_flag"][s]))


def _ostream_double(t):
    """A double as std::ostream prints it by default (%g, 6 digits)."""
    return "%g" % t


def state_file(dirname, t):
    """'results/<dir><t>' (io.hpp:112-113, 150-151)."""
    return f"results/{dirname}{_ostream_double(t)}"


def write_state(path, v):
    np.ascontiguousarray(v, np.float64).reshape(-1).tofile(path)


def read_state(path, n):
    v = np.fromfile(path, dtype=np.float64)
    if v.size != n:
        raise ValueError(f"{path} holds {v.size} values, expected {n}")
    return v


def _numbers(path):
    with open(path) as f:
        return f.read().split()


def read_lookup_tables(dirpath, name):
    """dict(p, T, wf_h2o_l, wf_co2_l) of magma `name` (T is None without a
    temperature.txt; the wf tables are then 1-D, else [n_T, n_p])."""
    import os
    tok = _numbers(os.path.join(dirpath, "pressure.txt"))
    n_p = int(tok[0])
    if n_p < 2 or len(tok) < 1 + n_p:
        raise ValueError("pressure.txt: a count >= 2 and that many values")
    p = np.array([float(x) for x in tok[1:1 + n_p]])
    T = None
    tpath = os.path.join(dirpath, "temperature.txt")
    if os.path.exists(tpath):
        tok = _numbers(tpath)
        n_T = int(tok[0])
        if n_T < 2 or len(tok) < 1 + n_T:
            raise ValueError("temperature.txt: a count >= 2 and that many values")
        T = np.array([float(x) for x in tok[1:1 + n_T]])
    n = n_p * (1 if T is None else T.size)
    out = dict(p=p, T=T)
    for key, fname in (("wf_h2o_l", "wf_l_h2o_onliq_"), ("wf_co2_l", "wf_l_co2_onliq_")):
        tok = _numbers(os.path.join(dirpath, fname + name + ".txt"))
        if len(tok) < n:
            raise ValueError(f"{fname}{name}.txt holds {len(tok)} values, expected {n}")
        v = np.array([float(x) for x in tok[:n]])
        out[key] = v if T is None else v.reshape(T.size, n_p)
    return out


def write_lookup_tables(dirpath, name, tables):
    """Writes tables (as read_lookup_tables returns them) in the same layout."""
    import os
    os.makedirs(dirpath, exist_ok=True)
    with open(os.path.join(dirpath, "pressure.txt"), "w") as f:
        f.write(f"{tables['p'].size}\n" + "".join(f"{x!r}\n" for x in tables["p"]))
    if tables.get("T") is not None:
        with open(os.path.join(dirpath, "temperature.txt"), "w") as f:
            f.write(f"{tables['T'].size}\n" + "".join(f"{x!r}\n" for x in tables["T"]))
    for key, fname in (("wf_h2o_l", "wf_l_h2o_onliq_"), ("wf_co2_l", "wf_l_co2_onliq_")):
        v = np.asarray(tables[key])
        with open(os.path.join(dirpath, fname + name + ".txt"), "w") as f:
            for row in v.reshape(-1, v.shape[-1]):
                f.write(" ".join(repr(float(x)) for x in row) + "\n")
