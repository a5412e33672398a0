"""Constants.set_constants -- src/constants.jl:6-13.  Module globals as in the reference."""
phys_c = 1.0
phys_a = 1.0
alpha = 1.0


def set_constants(inputs: dict) -> None:
    global phys_c, phys_a, alpha
    phys_c = float(inputs["PHYS_C"])
    phys_a = float(inputs["PHYS_A"])
    alpha = float(inputs["ALPHA"])
