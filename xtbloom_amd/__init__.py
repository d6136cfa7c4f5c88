"""xtbloom_amd: MI355X-native (gfx950) batched GFN2-xTB single points behind xTBloom's C ABI."""
__version__ = "0.1.0"
