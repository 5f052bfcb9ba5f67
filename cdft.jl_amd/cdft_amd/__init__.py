"""cdft_amd -- MI355X-native PC-SAFT classical DFT hot path (functional evaluation and
Euler-Lagrange Picard/Anderson fixed point) behind cDFT.jl's DFTSystem /
initialize_profiles / converge! API.  Compute runs in libcdft_hip.so (HIP, gfx950);
there is no CPU fallback."""
from .bulk import KB, NA, RGAS, PCSAFT
from .structure import (HardSphereMatrix, Steele, Structure, TwoPhase1DCart, TwoPhase2DLamCart, Uniform1DCart,
                        Uniform1DCyl, Uniform1DSphr, Uniform2DCart, Uniform3DCart)
from .system import (DFTOptions, DFTSystem, F_ideal, F_res, converge_, dFdrho_res, free_energy,
                     initialize_profiles, integrate, integration_weights, simpson)
from .methods import adsorption, interfacial_tension, surface_tension
from .ddft import ODEProblem, ddft_rhs

__all__ = ["PCSAFT", "Steele", "HardSphereMatrix", "Structure", "Uniform1DCart", "Uniform2DCart", "Uniform3DCart",
           "Uniform1DSphr", "Uniform1DCyl",
           "TwoPhase1DCart", "TwoPhase2DLamCart",
           "DFTOptions", "DFTSystem", "initialize_profiles", "converge_", "dFdrho_res", "F_res", "F_ideal",
           "free_energy", "simpson", "integrate", "integration_weights",
           "surface_tension", "interfacial_tension", "adsorption", "ddft_rhs", "ODEProblem", "NA", "KB", "RGAS"]
