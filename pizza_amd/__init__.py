"""pizza_amd — MI355X-native IMEX time step of pizza behind a C ABI.

``pizza_amd.runtime.Run`` drives one simulation through libpizza_amd.so;
``pizza_amd.config`` holds the namelist configuration.
"""
from .config import PzConfig, config_from_namelist, default_config  # noqa: F401
