"""cusoap_amd: MI355X-native SOAP descriptors (drop-in for TACC/cuSOAP's ``SOAP``)."""

from .soap import SOAP

__version__ = "0.1.0"
__all__ = ["SOAP"]
