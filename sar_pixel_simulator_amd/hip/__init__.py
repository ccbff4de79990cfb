"""HIP (MI355X) backend: simulations, batches and the serving loop."""
