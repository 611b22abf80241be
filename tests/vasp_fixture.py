"""Writes vasprun.xml files with known contents (public vasprun.xml layout) for
the training-data tests; the reference ships no vasprun files."""
import numpy as np


def _v(rows, name):
    body = "".join("    <v>" + " ".join(f"{x:16.8f}" for x in r) + " </v>\n" for r in rows)
    return f'   <varray name="{name}" >\n{body}   </varray>\n'


def _structure(lattice, frac, name=None):
    head = f'  <structure name="{name}" >\n' if name else "  <structure>\n"
    return (head + "   <crystal>\n" + _v(lattice, "basis")
            + f'    <i name="volume">{abs(np.linalg.det(lattice)):16.8f} </i>\n'
            + _v(np.linalg.inv(lattice).T, "rec_basis") + "   </crystal>\n"
            + _v(frac, "positions") + "  </structure>\n")


def write_vasprun(path, lattice, species, frac, eig, occ, dos_e, dos_tot, energies, steps=None):
    """eig, occ: [spin, k, band]; dos_tot: [spin, ne]; energies: e_fr_energy per
    ionic step; steps: optional list of (lattice, frac) per ionic step (default
    one step at the given structure)."""
    lattice = np.asarray(lattice, float)
    frac = np.asarray(frac, float)
    steps = steps or [(lattice, frac)]
    types = []
    for s in species:
        if s not in types:
            types.append(s)
    out = ['<?xml version="1.0" encoding="ISO-8859-1"?>\n<modeling>\n',
           ' <generator>\n  <i name="program" type="string">vasp </i>\n </generator>\n',
           ' <atominfo>\n', f'  <atoms>{len(species)} </atoms>\n', f'  <types>{len(types)} </types>\n',
           '  <array name="atoms" >\n   <dimension dim="1">ion</dimension>\n',
           '   <field type="string">element</field>\n   <field type="int">atomtype</field>\n   <set>\n']
    for s in species:
        out.append(f"    <rc><c>{s:2s}</c><c>{types.index(s) + 1:4d}</c></rc>\n")
    out.append("   </set>\n  </array>\n </atominfo>\n")
    out.append(_structure(steps[0][0], steps[0][1], "initialpos"))
    for i, (lat, fr) in enumerate(steps):
        out.append(" <calculation>\n")
        out.append('  <scstep>\n   <energy>\n    <i name="e_fr_energy">   1234.5 </i>\n   </energy>\n  </scstep>\n')
        out.append(_structure(np.asarray(lat, float), np.asarray(fr, float)))
        out.append("  <energy>\n")
        out.append(f'   <i name="e_fr_energy">{energies[i]:16.8f} </i>\n')
        out.append(f'   <i name="e_wo_entrp">{energies[i] + 0.001:16.8f} </i>\n')
        out.append(f'   <i name="e_0_energy">{energies[i] + 0.002:16.8f} </i>\n')
        out.append("  </energy>\n")
        if i == len(steps) - 1:
            out.append("  <eigenvalues>\n   <array>\n    <dimension dim=\"1\">band</dimension>\n"
                       "    <dimension dim=\"2\">kpoint</dimension>\n    <dimension dim=\"3\">spin</dimension>\n"
                       "    <field>eigene</field>\n    <field>occ</field>\n    <set>\n")
            for s in range(eig.shape[0]):
                out.append(f'     <set comment="spin {s + 1}">\n')
                for k in range(eig.shape[1]):
                    out.append(f'      <set comment="kpoint {k + 1}">\n')
                    for b in range(eig.shape[2]):
                        out.append(f"       <r> {eig[s, k, b]:12.4f} {occ[s, k, b]:8.4f} </r>\n")
                    out.append("      </set>\n")
                out.append("     </set>\n")
            out.append("    </set>\n   </array>\n  </eigenvalues>\n")
            out.append('  <dos>\n   <i name="efermi">      0.5 </i>\n   <total>\n    <array>\n'
                       "     <dimension dim=\"1\">gridpoints</dimension>\n     <dimension dim=\"2\">spin</dimension>\n"
                       "     <field>energy</field>\n     <field>total</field>\n     <field>integrated</field>\n"
                       "     <set>\n")
            for s in range(dos_tot.shape[0]):
                out.append(f'      <set comment="spin {s + 1}">\n')
                integ = np.cumsum(dos_tot[s])
                for e, d, c in zip(dos_e, dos_tot[s], integ):
                    out.append(f"       <r> {e:10.4f} {d:10.4f} {c:10.4f} </r>\n")
                out.append("      </set>\n")
            out.append("     </set>\n    </array>\n   </total>\n  </dos>\n")
        out.append(" </calculation>\n")
    out.append(_structure(steps[-1][0], steps[-1][1], "finalpos"))
    out.append("</modeling>\n")
    with open(path, "w") as f:
        f.write("".join(out))
