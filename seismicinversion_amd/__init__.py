"""MI355X-native Kennett recursive reflectivity (package bootstrap; filled in below)."""
