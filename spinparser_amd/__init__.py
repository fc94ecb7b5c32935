"""MI355X-native pf-FRG flow engine (SU(2) core), drop-in for SpinParser's FrgCore path."""
