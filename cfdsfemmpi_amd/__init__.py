"""MI355X-native stabilised SUPG/PSPG finite-element hot path (cfdsfemmpi drop-in)."""
