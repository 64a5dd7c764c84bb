"""Writes small capsomere templates in the reference's infiles format (the layout read by
src/initialize.cpp:46-128): three header lines, particles, edges, triangles, LJ attractions
and a (never read) repulsion section. Used to build known-answer systems for the tests."""


def write_template(path, beads, edges=(), faces=(), attractions=(), si_sigma=1e-9, si_mass=3.9434e-26):
    """beads: [(x, y, z, type, charge, diameter, mass)], edges: [(b1, b2, type, length)],
    faces: [(t1, t2, t3)], attractions: [(type1, type2)] ordered rows of lj_a."""
    L = [f"ks_vdwr  = 1.122462048", f"mass     = {si_mass!r}", f"sigma    = {si_sigma!r}", "",
         f"# Number of Particles = {len(beads)}", "", "index\tx\ty\tz\ttype\tcharge\tdiameter\tmass"]
    for i, (x, y, z, t, q, d, m) in enumerate(beads):
        L.append(f"{i}\t{float(x)!r}\t{float(y)!r}\t{float(z)!r}\t{int(t)}\t{float(q)!r}\t{float(d)!r}\t{float(m)!r}")
    L += ["", f"#     Number      of    Edges =     {len(edges)}", "", "index\te1\te2\ttype\tlength"]
    for i, (a, b, t, ln) in enumerate(edges):
        L.append(f"{i}\t{int(a)}\t{int(b)}\t{int(t)}\t{float(ln)!r}")
    L += ["", f"#     Number      of    Triangles   =     {len(faces)}", "", "index\tt1\tt2\tt3\ttype\tnormal"]
    for i, (a, b, c) in enumerate(faces):
        L.append(f"{i}\t{a}\t{b}\t{c}\t0\t1")
    L += ["", f"# Number of LJ Attractions = {len(attractions)}", "", "index  type1  type2  epsilon sigmahc"]
    for i, (a, b) in enumerate(attractions):
        L.append(f"{i}\t{a}\t{b}\t2\t1")
    L += ["", "# Number of LJ Repulsions = 1", "", "index  type1  type2  epsilon sigmahc", "0\t1\t1\t1\t1", ""]
    with open(path, "w") as f:
        f.write("\n".join(L))
    return path


def read_template(path):
    """The sections of a template in the same format, as write_template takes them: (beads,
    edges, faces, attractions, si_sigma, si_mass)."""
    tok = open(path).read().split()
    si_mass, si_sigma = float(tok[5]), float(tok[8])
    i = tok.index("Particles") + 2
    n = int(tok[i])
    i += 9
    beads = []
    for k in range(n):
        r = tok[i + 8 * k:i + 8 * k + 8]
        beads.append((float(r[1]), float(r[2]), float(r[3]), int(r[4]), float(r[5]), float(r[6]), float(r[7])))
    i = tok.index("Edges") + 2
    n = int(tok[i])
    i += 6
    edges = [(int(tok[i + 5 * k + 1]), int(tok[i + 5 * k + 2]), int(tok[i + 5 * k + 3]), float(tok[i + 5 * k + 4]))
             for k in range(n)]
    i = tok.index("Triangles") + 2
    n = int(tok[i])
    i += 7
    faces = [(int(tok[i + 6 * k + 1]), int(tok[i + 6 * k + 2]), int(tok[i + 6 * k + 3])) for k in range(n)]
    i = tok.index("Attractions") + 2
    n = int(tok[i])
    i += 6
    atts = [(int(tok[i + 5 * k + 1]), int(tok[i + 5 * k + 2])) for k in range(n)]
    return beads, edges, faces, atts, si_sigma, si_mass
