"""RegionalSeaLevel model choices (src/sealevel.jl:106-166, src/barystatic_sealevel.jl)."""


class LaterallyConstantSeaSurface:  pass   # src/sealevel.jl:122
class LaterallyVariableSeaSurface:  pass   # src/sealevel.jl:130
class NoSealevelLoad:               pass   # src/loads.jl:19
class InteractiveSealevelLoad:      pass   # src/loads.jl:26
class InternalBSLUpdate:            pass   # src/barystatic_sealevel.jl:60
class NoVolumeContribution:         pass   # src/barystatic_sealevel.jl:38
class GoelzerVolumeContribution:    pass   # src/barystatic_sealevel.jl:45
class NoDensityContribution:        pass   # src/barystatic_sealevel.jl:70
class GoelzerDensityContribution:   pass   # src/barystatic_sealevel.jl:77
class NoAdjustmentContribution:     pass   # src/barystatic_sealevel.jl:94
class GoelzerAdjustmentContribution: pass  # src/barystatic_sealevel.jl:101


class ConstantBSL:
    """Barystatic sea level held at its reference value z (src/barystatic_sealevel.jl:100-110,
    215-218). The reference builds the ocean-surface table of ReferenceBSL by download; a constant
    BSL never reads it, so only z is kept."""
    def __init__(self, z=0.0):
        self.z = float(z)


class RegionalSeaLevel:
    def __init__(self, surface=None, load=None, bsl=None, update_bsl=None, volume_contribution=None,
                 density_contribution=None, adjustment_contribution=None):
        self.surface = surface or LaterallyConstantSeaSurface()
        self.load = load or NoSealevelLoad()
        self.bsl = bsl or ConstantBSL()
        self.update_bsl = update_bsl or InternalBSLUpdate()
        self.volume_contribution = volume_contribution or GoelzerVolumeContribution()
        self.density_contribution = density_contribution or GoelzerDensityContribution()
        self.adjustment_contribution = adjustment_contribution or NoAdjustmentContribution()
        if not isinstance(self.bsl, ConstantBSL):
            raise NotImplementedError("only ConstantBSL is supported by this engine")
