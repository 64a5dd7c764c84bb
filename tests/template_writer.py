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
