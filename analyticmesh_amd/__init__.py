"""MI355X-native analytic marching (exact meshing of ReLU-MLP level sets)."""
