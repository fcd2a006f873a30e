"""MI355X-native wall-corrected RPY blob mobility (RigidMultiblobsWall hot path).

Submodules
----------
mobility   drop-in products with the reference signatures (numpy in/out)
wall_rpy   device-level products on torch ROCm tensors
workloads  synthetic benchmark configurations
"""
__version__ = "0.1.0"
