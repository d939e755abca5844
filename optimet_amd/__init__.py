"""MI355X-native backend for OPTIMET's multiple-scattering hot path."""
