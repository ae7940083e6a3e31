"""zorro_amd: MI355X-native drift registration for Zorro movies."""
