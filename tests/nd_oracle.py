"""Oracle for the temporal reuse of a nested-dissection separator tree (test infrastructure only):
plain-loop restatements of the reference's rules on CParent / CMember arrays
(ParallelNestedDissection/persistent_separator_tree.hh:63-72 lca, :79-113 dirty_subtrees,
:115-138 subtree nodes and vertices) and of what makes a forest a valid separator tree for a
graph (every edge joins two components on one root path)."""


def depths(parent):
    d = [0] * len(parent)
    for c in range(len(parent) - 1, -1, -1):          # parents have larger ids (postorder)
        d[c] = 0 if parent[c] < 0 else d[parent[c]] + 1
    return d


def lca(parent, depth, a, b):
    while depth[a] > depth[b]:
        a = parent[a]
    while depth[b] > depth[a]:
        b = parent[b]
    while a != b and a >= 0 and b >= 0:
        a, b = parent[a], parent[b]
    return a if a == b else -1


def edges_of(p, i):
    """Undirected edges (u < v) of a CSC graph of both triangles."""
    out = set()
    for v in range(len(p) - 1):
        for k in range(p[v], p[v + 1]):
            u = int(i[k])
            if u != v:
                out.add((min(u, v), max(u, v)))
    return out


def is_separator_tree(parent, member, edges):
    d = depths(parent)
    for (u, v) in edges:
        a, b = member[u], member[v]
        r = lca(parent, d, a, b)
        if r != a and r != b:
            return False
    return True


def dirty_subtrees(parent, member, edges):
    """(full_rebuild, outermost violated components in increasing id order)."""
    d = depths(parent)
    roots = [c for c in range(len(parent)) if parent[c] < 0]
    dirty = [False] * len(parent)
    for (u, v) in sorted(edges):
        a, b = member[v], member[u]
        r = lca(parent, d, a, b)
        if r == a or r == b:
            continue
        if r < 0 or (parent[r] < 0 and len(roots) == 1):
            return True, []
        dirty[r] = True
    out = []
    for c in range(len(parent)):
        if dirty[c]:
            p = parent[c]
            while p >= 0 and not dirty[p]:
                p = parent[p]
            if p < 0:
                out.append(c)
    return False, out


def subtree_vertices(parent, member, root):
    inside = set()
    for c in range(len(parent)):
        x = c
        while x >= 0 and x != root:
            x = parent[x]
        if x == root:
            inside.add(c)
    return [v for v in range(len(member)) if member[v] in inside]
