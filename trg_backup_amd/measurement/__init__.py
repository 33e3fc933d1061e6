"""Drop-ins for the initial-tensor builders of the reference's measurement drivers."""
