"""orbit -- Python 3 layer of the MI355X build of py-orbit's tracking + space-charge path.

Subpackages keep the reference's names (py/orbit/*): lattice, teapot, teapot_base,
space_charge (sc2p5d, sc3d), bunch_generators, utils.
"""
