"""acoular_amd -- Acoular's frequency-domain beamforming path, rebuilt for AMD MI355X (gfx950).

The public names mirror Acoular's (``PowerSpectra``, ``PowerSpectraImport``, ``SteeringVector``,
``BeamformerBase``, ``RectGrid``, ``MicGeom``, ``Environment``, ``L_p``, and the ``fastFuncs``
kernels) so that code using that path can switch imports.  All compute runs in hand-written
HIP kernels in ``libabf.so`` (C ABI in ``include/abf.h``); PyTorch supplies device memory and
streams.
"""

from .environments import Environment, dist_mat
from .fastFuncs import beamformerFreq, calcCSM, calcTransfer
from .fbeamform import (
    BeamformerBase,
    BeamformerCapon,
    BeamformerClean,
    BeamformerCleansc,
    BeamformerDamas,
    BeamformerEig,
    BeamformerFunctional,
    BeamformerMusic,
    BeamformerOrth,
    LazyBfResult,
    L_p,
    PointSpreadFunction,
    SteeringVector,
    beamform_device,
    beamform_eig_device,
    beamform_f64_device,
    clean_device,
    cleansc_device,
    damas_device,
    integrate,
    psf_device,
)
from .fprocess import IRFFT, RFFT, AutoPowerSpectra, CrossPowerSpectra
from .grids import (
    CircSector,
    ConvexSector,
    Grid,
    ImportGrid,
    LineGrid,
    MergeGrid,
    MultiSector,
    PolySector,
    RectGrid,
    RectGrid3D,
    RectSector,
    RectSector3D,
    Sector,
    SingleSector,
)
from .microphones import MicGeom
from .signals import NoiseGenerator, SignalGenerator, SineGenerator, WNoiseGenerator
from .sources import PointSource, SourceMixer, TimeSamples, UncorrelatedNoiseSource
from .spectra import BaseSpectra, PowerSpectra, PowerSpectraImport

__all__ = [
    'CircSector',
    'LineGrid',
    'MergeGrid',
    'ConvexSector',
    'MultiSector',
    'PolySector',
    'RectSector',
    'RectSector3D',
    'Sector',
    'SingleSector',
    'RFFT',
    'IRFFT',
    'AutoPowerSpectra',
    'CrossPowerSpectra',
    'BaseSpectra',
    'BeamformerBase',
    'BeamformerCapon',
    'BeamformerClean',
    'BeamformerCleansc',
    'BeamformerDamas',
    'BeamformerEig',
    'BeamformerFunctional',
    'BeamformerMusic',
    'BeamformerOrth',
    'Environment',
    'Grid',
    'ImportGrid',
    'L_p',
    'LazyBfResult',
    'MicGeom',
    'PowerSpectra',
    'NoiseGenerator',
    'PointSource',
    'PointSpreadFunction',
    'PowerSpectraImport',
    'RectGrid',
    'RectGrid3D',
    'SignalGenerator',
    'SineGenerator',
    'SourceMixer',
    'SteeringVector',
    'TimeSamples',
    'UncorrelatedNoiseSource',
    'WNoiseGenerator',
    'beamform_device',
    'beamform_eig_device',
    'beamform_f64_device',
    'clean_device',
    'cleansc_device',
    'damas_device',
    'beamformerFreq',
    'calcCSM',
    'calcTransfer',
    'dist_mat',
    'integrate',
    'psf_device',
]
