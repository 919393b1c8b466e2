"""POT3D's driver side, restated for the tests, the bench and the warm-start
training: meshes (genmesh), the synthetic Br maps and the named workloads.
The solver library (pot3d_amd) takes what these produce as its inputs."""
